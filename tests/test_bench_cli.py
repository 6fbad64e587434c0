"""CPU checks of bench.py's command line (no GPU needed)."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), *args], capture_output=True,
                          text=True, timeout=120, env=e)


def test_more_than_one_gpu_exits_with_a_message():
    p = _run("--gpus", "2", "--steps", "1", "--warmup", "0")
    assert p.returncode == 2
    assert "measures one GPU" in p.stderr
    assert p.stdout == ""


def test_world_size_above_one_exits_with_a_message():
    p = _run("--steps", "1", env={"WORLD_SIZE": "4"})
    assert p.returncode == 2
    assert "WORLD_SIZE 4" in p.stderr


def test_roofline_profiles_carry_traffic():
    """The committed PMC summaries the bench line quotes parse to a traffic
    figure no smaller than the kernel's algorithmic bytes."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(REPO, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    n = 240 * 240 * 140
    for name, alg in (("stencil_fwd_c2_pmc.txt", 40.0 * n), ("gs_update_c2_pmc.txt", None),
                      ("gs_dots_c2_pmc.txt", None)):
        t = bench._traffic_from_profile(name)
        assert t is not None and t > 0, name
        if alg is not None:
            assert alg <= t < 1.5 * alg, (name, t, alg)
