"""GPU FGMRES(20) against the CPU oracle of FGMRES.m on the same operator.

Bar: the GPU solve reaches ||b - A x|| / ||b|| < tol measured with the oracle
operator in FP64, in an iteration count within 2% (+2) of the oracle's; the
quasi-residual history agrees to 1e-6 relative over the first 40 steps.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _setup(cfg):
    from helm3d.models import extend, wavenumber, vec, point_source
    vp = extend(cfg.v, cfg.npml)
    wn = vec(wavenumber(vp, cfg.freq))
    b = point_source(cfg, cfg.sources[0])
    return wn, b


@pytest.mark.parametrize("n", [(5, 4, 4), (3, 1, 2), (1, 1, 1)])
@pytest.mark.parametrize("mode", [1, 2])
def test_solve_tiny_grids(n, mode):
    """FGMRES(20) on grids of 1 to 80 points: one reduction block, vectors
    shorter than a block's threads, and (3 x 1 x 2, 1 x 1 x 1) a Krylov space
    exhausted before the tolerance is met (the step after is a breakdown)."""
    import torch
    from helm3d import Helm3d
    N = int(np.prod(n))
    rng = np.random.default_rng(5)
    wn = rng.uniform(2e-4, 2e-3, N)
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    npml = np.array([[1, 1, 1], [1, 1, 1]])
    h = (10.0, 10.0, 10.0)
    variant = "forw" if mode == 1 else "adj"
    print(f"tiny solve n={n} mode={mode}", flush=True)
    xo, info_o = O.fgmres(variant, wn, h, n, npml, b, tol=1e-6, maxinnerit=20)
    dev = torch.device("cuda:0")
    H = Helm3d(n, h, npml, torch.from_numpy(wn).to(dev))
    x, info = H.solve(torch.from_numpy(b).to(dev), mode=mode, tol=1e-6, maxinnerit=20)
    x = x.cpu().numpy()
    assert np.isfinite(x).all()
    relres = np.linalg.norm(b - O.hmvp(variant, wn, h, n, npml, x)) / np.linalg.norm(b)
    print("gpu iters", info.iters, "oracle iters", info_o["iters"], "relres", relres, flush=True)
    assert info.converged and relres < 1e-6
    assert abs(info.iters - info_o["iters"]) <= 0.02 * info_o["iters"] + 2
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-3


@pytest.mark.parametrize("n", [(5, 4, 4), (3, 1, 2), (1, 1, 1)])
@pytest.mark.parametrize("mode", [1, 2])
def test_bicgstab_and_mixed_tiny_grids(n, mode):
    """FP64 BiCGStab and the mixed-precision solve (complex64 BiCGStab to a
    tolerance, complex64 GMRES, and the fixed-length BiCGStab replayed from a
    graph) on grids of 1 to 80 points.  On 1 x 1 x 1 one iteration solves the
    system, so the fixed-length inner solve runs on with zero or rounding-
    sized residuals, where its scalar kernels' zero-denominator guards give
    zero steps."""
    import torch
    from helm3d import Helm3d
    N = int(np.prod(n))
    rng = np.random.default_rng(5)
    wn = rng.uniform(2e-4, 2e-3, N)
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    npml = np.array([[1, 1, 1], [1, 1, 1]])
    h = (10.0, 10.0, 10.0)
    variant = "forw" if mode == 1 else "adj"
    dev = torch.device("cuda:0")
    H = Helm3d(n, h, npml, torch.from_numpy(wn).to(dev))
    bt = torch.from_numpy(b).to(dev)
    xf, _ = O.fgmres(variant, wn, h, n, npml, b, tol=1e-12, maxinnerit=20)

    def check(x, info, what):
        x = x.cpu().numpy()
        rel = np.linalg.norm(b - O.hmvp(variant, wn, h, n, npml, x)) / np.linalg.norm(b)
        err = np.linalg.norm(x - xf) / np.linalg.norm(xf)
        print(f"  {what}: iters={info.iters} relres={info.relres:.1e} (oracle {rel:.1e}) err={err:.1e}", flush=True)
        assert np.isfinite(x).all() and info.converged
        assert rel < 1e-6 and abs(rel - info.relres) < 1e-9
        assert err < 1e-3

    print(f"tiny bicgstab/mixed n={n} mode={mode}", flush=True)
    with np.errstate(all="ignore"):
        _, it_o, _ = O.bicgstab(variant, wn, h, n, npml, b, tol=1e-6, maxit=200)
    x, info = H.solve_bicgstab(bt, mode=mode, tol=1e-6, maxit=200)
    check(x, info, f"bicgstab fp64 (oracle {it_o} iterations)")
    assert 0.75 * it_o <= info.iters // 2 <= 1.33 * it_o + 2
    for inner, itol in (("bicgstab", 1e-2), ("gmres", 1e-2), ("bicgstab", 0.0)):
        x, info = H.solve_mixed(bt, mode=mode, tol=1e-6, maxit=200, inner=inner, inner_tol=itol, inner_maxit=10)
        check(x, info, f"mixed {inner} inner_tol={itol}")


@pytest.mark.parametrize("mode", [1, 2])
def test_solve_matches_oracle(mode):
    import torch
    from helm3d import Helm3d
    from helm3d.models import config_small
    cfg = config_small()
    wn, b = _setup(cfg)
    npml = cfg.npml6
    h = (cfg.h,) * 3
    variant = "forw" if mode == 1 else "adj"
    print(f"solve mode={mode} nt={cfg.nt}", flush=True)
    xo, info_o = O.fgmres(variant, wn, h, cfg.nt, np.array(npml), b, tol=1e-6, maxinnerit=20)
    dev = torch.device("cuda:0")
    H = Helm3d(cfg.nt, h, npml, torch.from_numpy(wn).to(dev))
    x, info = H.solve(torch.from_numpy(b).to(dev), mode=mode, tol=1e-6, maxinnerit=20, return_hst=True)
    x = x.cpu().numpy()
    r = b - O.hmvp(variant, wn, h, cfg.nt, np.array(npml), x)
    relres = np.linalg.norm(r) / np.linalg.norm(b)
    print("gpu iters", info.iters, "oracle iters", info_o["iters"], "relres", relres, info.relres)
    assert info.converged and relres < 1e-6
    assert abs(relres - info.relres) <= 1e-9
    assert abs(info.iters - info_o["iters"]) <= 0.02 * info_o["iters"] + 2
    k = min(40, len(info.hst), len(info_o["hst"]))
    assert np.allclose(info.hst[:k], info_o["hst"][:k], rtol=1e-6, atol=0)
    # both solutions solve the same system to the same tolerance
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-3


def test_solve_zero_rhs_and_x0():
    import torch
    from helm3d import Helm3d
    from helm3d.models import config_small
    cfg = config_small(n=(16, 16, 12), npml=4)
    wn, b = _setup(cfg)
    dev = torch.device("cuda:0")
    H = Helm3d(cfg.nt, (cfg.h,) * 3, cfg.npml6, torch.from_numpy(wn).to(dev))
    z = torch.zeros(b.size, dtype=torch.complex128, device=dev)
    x, info = H.solve(z)
    assert info.converged and float(x.abs().max()) == 0.0
    bt = torch.from_numpy(b).to(dev)
    h3, p6 = (cfg.h,) * 3, np.array(cfg.npml6)
    x1, i1 = H.solve(bt, tol=1e-3)
    # warm start: FGMRES.m:57-58,67 measures tol relative to ||b - A x0||
    x2, i2 = H.solve(bt, tol=1e-6, x0=x1)
    x1n = x1.cpu().numpy()
    xo2, io2 = O.fgmres("forw", wn, h3, cfg.nt, p6, b, x0=x1n, tol=1e-6, maxinnerit=20)
    r0 = b - O.hmvp("forw", wn, h3, cfg.nt, p6, x1n)
    r2 = b - O.hmvp("forw", wn, h3, cfg.nt, p6, x2.cpu().numpy())
    assert abs(np.linalg.norm(r2) / np.linalg.norm(r0) - i2.relres) <= 1e-9
    assert i2.converged and i2.relres < 1e-6
    assert abs(i2.iters - io2["iters"]) <= 0.02 * io2["iters"] + 2


def test_bicgstab_fp64_matches_oracle():
    import torch
    from helm3d import Helm3d
    from helm3d.models import config_small
    cfg = config_small(n=(24, 24, 20), npml=6)
    wn, b = _setup(cfg)
    h3, p6 = (cfg.h,) * 3, np.array(cfg.npml6)
    xo, it_o, hist = O.bicgstab("forw", wn, h3, cfg.nt, p6, b, tol=1e-6)
    dev = torch.device("cuda:0")
    H = Helm3d(cfg.nt, h3, cfg.npml6, torch.from_numpy(wn).to(dev))
    x, info = H.solve_bicgstab(torch.from_numpy(b).to(dev), tol=1e-6)
    print("bicgstab gpu matvecs", info.iters, "oracle iters", it_o, info.relres)
    assert info.converged
    # BiCGStab on this indefinite operator amplifies rounding differences of
    # the operator (1e-16 relative) into different iterate paths, so the step
    # counts agree only roughly; the solutions must agree to the tolerance.
    assert 0.75 * it_o <= info.iters // 2 <= 1.33 * it_o + 2
    xg = x.cpu().numpy()
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-4
    r = b - O.hmvp("forw", wn, h3, cfg.nt, p6, xg)
    assert abs(np.linalg.norm(r) / np.linalg.norm(b) - info.relres) < 1e-9
    assert info.relres < 1e-5


@pytest.mark.parametrize("inner", ["bicgstab", "gmres"])
def test_mixed_precision_solve(inner):
    import torch
    from helm3d import Helm3d
    from helm3d.models import config_small
    cfg = config_small()
    wn, b = _setup(cfg)
    dev = torch.device("cuda:0")
    H = Helm3d(cfg.nt, (cfg.h,) * 3, cfg.npml6, torch.from_numpy(wn).to(dev))
    bt = torch.from_numpy(b).to(dev)
    x64, i64 = H.solve(bt, tol=1e-6)
    xm, im = H.solve_mixed(bt, tol=1e-6, inner_tol=1e-2, inner=inner)
    print(inner, "plain", i64.iters, "mixed outer", im.iters, "inner", im.inner_iters, im.relres)
    r = b - O.hmvp("forw", wn, (cfg.h,) * 3, cfg.nt, np.array(cfg.npml6), xm.cpu().numpy())
    assert im.converged and np.linalg.norm(r) / np.linalg.norm(b) < 1e-6
    # per-solve error against the complex128 solution, both at residual 1e-6
    err = float(torch.linalg.norm(xm - x64) / torch.linalg.norm(x64))
    print("error vs c128 solve", err)
    assert err < 1e-3
    # adjoint system
    xa, ia = H.solve_mixed(bt, mode=2, tol=1e-6, inner=inner)
    ra = b - O.hmvp("adj", wn, (cfg.h,) * 3, cfg.nt, np.array(cfg.npml6), xa.cpu().numpy())
    assert ia.converged and np.linalg.norm(ra) / np.linalg.norm(b) < 1e-6


@pytest.mark.parametrize("m", [20, 5])
def test_captured_arnoldi_cycles_match_eager(m):
    """On small grids a whole Arnoldi cycle is replayed from a captured HIP
    graph: every cycle of every solve (eager residuals and combines between
    the replays, forward and adjoint re-captures) gives the bits of the
    step-by-step launches."""
    import torch
    from helm3d import Helm3d
    from helm3d.models import config_small
    cfg = config_small(n=(30, 28, 24), npml=6)
    wn, b = _setup(cfg)
    dev = torch.device("cuda:0")
    H = Helm3d(cfg.nt, (cfg.h,) * 3, cfg.npml6, torch.from_numpy(wn).to(dev))
    bt = torch.from_numpy(b).to(dev)
    ref = {}
    H.set_graphs(False)
    for mode in (1, 2):
        ref[mode] = H.solve(bt, mode=mode, tol=1e-8, maxinnerit=m)
    H.set_graphs(True)
    for rep in range(2):
        for mode in (1, 2):
            x, info = H.solve(bt, mode=mode, tol=1e-8, maxinnerit=m)
            xe, ie = ref[mode]
            print("mode", mode, "rep", rep, "iters", info.iters, ie.iters, "cycles", info.cycles, flush=True)
            assert info.iters == ie.iters and info.cycles == ie.cycles and torch.equal(x, xe)
            H.apply(bt)  # eager work between replays
            torch.cuda.synchronize()
    assert ref[1][1].cycles > 2  # several replays within one solve


def test_multigrid_graphs_match_eager():
    """Multigrid smoothers replay captured cycles on the coarse levels: the
    solution equals the step-by-step one bit for bit."""
    import torch
    from helm3d.mg import MLGMRES
    from helm3d.models import config_small, point_source, vec
    cfg = config_small(n=(40, 36, 30), npml=8)
    dev = torch.device("cuda:0")
    M = MLGMRES(vec(cfg.v), cfg.n_phys, cfg.h, cfg.npml, cfg.freq, nlevels=3, device=dev, autotune=False)
    b = torch.from_numpy(point_source(cfg, cfg.sources[0])).to(dev)
    out = {}
    for g in (False, True, True):
        for H in M.ops + M.smoothers:
            H.set_graphs(g)
        x, info = M.solve(b, tol=1e-8, precision="mixed")
        if g in out:
            assert torch.equal(x, out[g][0])
        out[g] = (x, info)
    assert out[True][1].iters == out[False][1].iters and torch.equal(out[True][0], out[False][0])


def test_mixed_fixed_inner_graph_replays_match_eager():
    """Fixed-length complex64 inner solves replayed from a captured HIP graph
    give bit-identical solutions to eager launches, over repeated replays with
    eager work in between."""
    import torch
    from helm3d import Helm3d
    from helm3d.models import config_small
    cfg = config_small(n=(24, 24, 20), npml=6)
    wn, b = _setup(cfg)
    dev = torch.device("cuda:0")
    H = Helm3d(cfg.nt, (cfg.h,) * 3, cfg.npml6, torch.from_numpy(wn).to(dev))
    bt = torch.from_numpy(b).to(dev)
    H.set_graphs(False)
    xe, ie = H.solve_mixed(bt, tol=1e-6, inner_tol=0.0, inner_maxit=10)
    H.set_graphs(True)
    for rep in range(3):
        xg, ig = H.solve_mixed(bt, tol=1e-6, inner_tol=0.0, inner_maxit=10)
        assert ig.iters == ie.iters and torch.equal(xg, xe), rep
        H.apply(bt)  # eager work between replays
        torch.cuda.synchronize()
    r = b - O.hmvp("forw", wn, (cfg.h,) * 3, cfg.nt, np.array(cfg.npml6), xg.cpu().numpy())
    assert ig.converged and np.linalg.norm(r) / np.linalg.norm(b) < 1e-6
    # the adjoint re-captures (different operator variant)
    xa, ia = H.solve_mixed(bt, mode=2, tol=1e-6, inner_tol=0.0, inner_maxit=10)
    ra = b - O.hmvp("adj", wn, (cfg.h,) * 3, cfg.nt, np.array(cfg.npml6), xa.cpu().numpy())
    assert ia.converged and np.linalg.norm(ra) / np.linalg.norm(b) < 1e-6


@pytest.mark.parametrize("m,maxit", [(20, 2), (7, 3), (1, 5)])
def test_solve_stops_at_maxit_like_oracle(m, maxit):
    """FGMRES.m's loop guard floor(it/m) < maxit (it counted from 1) with the
    Arnoldi steps enqueued ahead of the host: the device stops after the
    oracle's number of steps with its iterate and quasi-residual history (a
    step enqueued past the stopping point is discarded)."""
    import torch
    from helm3d import Helm3d
    from helm3d.models import config_small
    cfg = config_small()
    wn, b = _setup(cfg)
    npml, h = cfg.npml6, (cfg.h,) * 3
    print(f"maxit stop m={m} maxit={maxit}", flush=True)
    xo, io = O.fgmres("forw", wn, h, cfg.nt, np.array(npml), b, tol=1e-12, maxinnerit=m, maxit=maxit)
    H = Helm3d(cfg.nt, h, npml, torch.from_numpy(wn).cuda())
    x, info = H.solve(torch.from_numpy(b).cuda(), tol=1e-12, maxinnerit=m, maxit=maxit, return_hst=True)
    print("gpu", info.iters, info.cycles, info.relres, "oracle", io["iters"], io["relres"], flush=True)
    assert not info.converged
    assert info.iters == io["iters"] and info.cycles <= maxit
    assert abs(info.relres - io["relres"]) <= 1e-8 * io["relres"]
    k = min(len(info.hst), len(io["hst"]))
    assert k >= info.iters
    assert np.allclose(info.hst[:k], io["hst"][:k], rtol=1e-8, atol=0)
    x = x.cpu().numpy()
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-8
