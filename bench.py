#!/usr/bin/env python3
"""Benchmark of the 3D Helmholtz hot path on one MI355X.

Workload (one "step"): one right-hand side of the plain complex128 method
solved to ||b - A x|| / ||b|| < 1e-6 with restarted GMRES(20) (FGMRES.m,
identity preconditioner) on the 27-point stencil with PML:
200x200x100 physical grid + 20-point PML (240x240x140 = 8,064,000 points),
h = 10 m, 6 Hz, velocity 1500->4000 m/s linear in depth with a seeded smooth
+-10% perturbation, unit point source at (100,100,5).  Synthetic data, inputs
resident in HBM before the timed region.

value = solved RHS/s over the timed steps (nothing instrumented inside them).
The JSON line also carries
  roofline      the kernel class that takes most of the solve's time (the
                classical Gram-Schmidt update w -= V h with ||w||^2,
                16 (K+2) B/pt for K basis vectors): algorithmic bytes over
                its launch durations, timed with HIP events on the library's
                stream over a second pass of the same K solves (the rate of
                that instrumented pass is reported beside value); traffic
                from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary in
                profiles/;
  stencil_roofline  the same for the stencil (40 B/pt matvec, 56 B/pt
                fused residual);
  mixed_batched the batched mixed-precision solve of the same model with 8
                sources on a surface line (FP64 flexible GMRES, 20 complex64
                BiCGStab iterations per outer step, all 8 right-hand sides in
                one launch per kernel type): RHS/s, ratio to the plain path,
                and each solution's error against the complex128 solution of
                the same source;
  multigrid, multigrid_sources, batched_uneven  the other solver forms;
  fwi_grads_per_min  least-squares FWI gradients w.r.t. squared slowness on
                256x256x128 + 20-pt PML at 12.5 m (16 sources, 64x64
                receivers, 4 and 6 Hz, solves to 1e-6): one full gradient
                (32 forward + 32 adjoint solves) timed;
  cpu_baseline  the reference's own C stencil (do_Hmvp compiled from the
                reference header in place, OpenMP, one thread per usable
                host core, bound close to cores) driving a C restatement of
                FGMRES.m: the whole 84^3 (64^3 + 10-pt PML) solve timed end to
                end beside the same solve on the GPU, the 240x240x140 matvec,
                and a bounded sample of the 240x240x140 solve.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "slim-release-apps_amd")
HBM_PEAK_GBS = 8000.0
METRIC = "solved RHS/sec at 1e-6 FP64 residual + FWI grads/min; matvec % HBM peak"


def host_cpu_info():
    """The host CPUs this process may use: the affinity mask, the cgroup CPU
    quota (cpu.max) and the OpenMP share the environment sets
    (OMP_NUM_THREADS); the CPU baseline uses all of them, as many threads as
    the smallest of the three allows."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = nproc
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    omp_n = int(omp) if omp and omp.isdigit() and int(omp) > 0 else None
    threads = affinity
    if quota is not None:
        threads = min(threads, max(1, int(quota)))
    if omp_n is not None:
        threads = min(threads, omp_n)
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": nproc, "affinity": affinity, "cgroup_quota_cpus": quota, "OMP_NUM_THREADS": omp,
            "threads": max(1, threads), "model": model}


def _cpu_env():
    """Environment of the CPU-baseline child process: OpenMP threads bound one
    per core, packed.  Set only there: bound at start-up, an OpenMP runtime
    pins the initial thread to one core, and every thread it creates later
    (the GPU legs' host threads, the affinity query) would inherit that one
    core."""
    env = dict(os.environ)
    env["OMP_PROC_BIND"] = "close"
    env["OMP_PLACES"] = "cores"
    return env


def _alg_bytes(kind, iters, cycles, m, N, stencil_launches):
    """Algorithmic HBM bytes of one solve per kernel class (complex128)."""
    if kind == "stencil":
        resid = cycles
        return (stencil_launches - resid) * 40.0 * N + resid * 56.0 * N
    # Arnoldi step j (1-based within its cycle) reads j basis vectors plus w
    steps, left = [], iters
    while left > 0:
        k = min(m, left)
        steps.append(k)
        left -= k
    if kind == "dots":
        return sum(16.0 * (j + 1) * N for k in steps for j in range(1, k + 1))
    if kind == "update":
        return sum(16.0 * (j + 2) * N for k in steps for j in range(1, k + 1))
    if kind == "combine":
        return sum(16.0 * (k + 2) * N for k in steps)
    return 0.0


def _traffic_from_profile(name):
    """HBM bytes per launch measured with rocprofv3 PMC counters (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), as recorded in
    profiles/<name>; None when absent."""
    path = os.path.join(REPO, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def _load_models():
    """helm3d/models.py (numpy / scipy only) without importing the package,
    which would import torch and its OpenMP runtime."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("_bench_models", os.path.join(PKG, "helm3d", "models.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = mod  # dataclasses look their module up while the file runs
    spec.loader.exec_module(mod)
    return mod


def cpu_mode(args):
    """CPU part of the baseline (run as a child process with _cpu_env()): the
    reference's C stencil (do_Hmvp compiled from its header in place, gcc -O3
    -fopenmp, else the oracle's C restatement) driving the C restatement of
    FGMRES.m on the host cores:
      * the whole C1 solve (84^3, 5 Hz, one impulse, FGMRES(20) to 1e-6);
      * the C2 matvec (2 warm-ups, median of 10);
      * a bounded sample of the C2 solve (sample_iters Arnoldi steps).
    Prints one JSON line; touches no GPU."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import numpy as np
    import oracle as O
    M = _load_models()
    host = host_cpu_info()
    th = host["threads"]
    use_ref = O.ref_available()
    c1 = M.config_c1()
    wn1 = M.vec(M.wavenumber(M.extend(c1.v, c1.npml), c1.freq))
    b1 = M.point_source(c1, c1.sources[0])
    t0 = time.perf_counter()
    _, i1 = O.fgmres("forw", wn1, (c1.h,) * 3, c1.nt, np.array(c1.npml6), b1, tol=1e-6, maxinnerit=20, nthreads=th,
                     use_ref=use_ref)
    c1_s = time.perf_counter() - t0
    c2 = M.config_c2()
    wn2 = M.vec(M.wavenumber(M.extend(c2.v, c2.npml), c2.freq))
    h2, p2 = (c2.h,) * 3, np.array(c2.npml6)
    x = np.random.default_rng(0).standard_normal(wn2.size) + 0j
    for _ in range(2):
        O.hmvp("forw", wn2, h2, c2.nt, p2, x, nthreads=th, use_ref=use_ref)
    ts = []
    for _ in range(10):
        t1 = time.perf_counter()
        O.hmvp("forw", wn2, h2, c2.nt, p2, x, nthreads=th, use_ref=use_ref)
        ts.append(time.perf_counter() - t1)
    mv = float(np.median(ts))
    b2 = M.point_source(c2, c2.sources[0])
    t0 = time.perf_counter()
    _, i2 = O.fgmres("forw", wn2, h2, c2.nt, p2, b2, tol=1e-6, maxinnerit=20, nthreads=th, use_ref=use_ref,
                     max_iters_cap=args.cpu_sample_iters)
    c2_s = time.perf_counter() - t0
    print(json.dumps({"mode": "cpu-baseline", "host": host, "threads": th, "use_ref": use_ref,
                      "omp": {k: os.environ.get(k) for k in ("OMP_PROC_BIND", "OMP_PLACES", "OMP_NUM_THREADS")},
                      "c1_seconds": c1_s, "c1_iters": i1["iters"], "c1_relres": i1["relres"],
                      "c2_matvec_s": mv, "c2_points": int(wn2.size),
                      "c2_sample_seconds": c2_s, "c2_sample_iters": i2["iters"]}))


def cpu_baseline(gpu_c2_iters, sample_iters):
    """The CPU baseline beside the GPU: the CPU part in a child process
    (cpu_mode), the same C1 solve on the GPU here."""
    import subprocess
    import numpy as np
    import torch
    from helm3d import Helm3d
    from helm3d.models import config_c1, extend, point_source, vec, wavenumber
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--mode", "cpu-baseline", "--cpu-sample-iters",
                        str(sample_iters)], env=_cpu_env(), capture_output=True, text=True, timeout=1200)
    if r.returncode != 0:
        raise RuntimeError(f"CPU baseline child failed ({r.returncode}): {r.stderr[-2000:]}")
    cpu = json.loads(r.stdout.strip().splitlines()[-1])
    c1 = config_c1()
    wn1 = vec(wavenumber(extend(c1.v, c1.npml), c1.freq))
    dev = torch.device("cuda:0")
    H1 = Helm3d(c1.nt, (c1.h,) * 3, c1.npml6, torch.from_numpy(wn1).to(dev))
    bt1 = torch.from_numpy(point_source(c1, c1.sources[0])).to(dev)
    H1.solve(bt1, tol=1e-6)  # warm-up (workspace)
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        xg, ig = H1.solve(bt1, tol=1e-6)
    torch.cuda.synchronize()
    gpu_s = (time.perf_counter() - t0) / reps
    del H1, xg
    th, c1_s, N2 = cpu["threads"], cpu["c1_seconds"], cpu["c2_points"]
    return {
        "value": 1.0 / c1_s,
        "unit": "RHS/s",
        "cores": th,
        "kind": "reference" if cpu["use_ref"] else "port",
        "sample": (f"the whole C1 solve (64^3 + 10-pt PML = 84^3, 5 Hz, one impulse at (32,32,5), FGMRES(20) to "
                   f"1e-6): {cpu['c1_iters']} iterations in {c1_s:.1f} s on {th} OpenMP threads (OMP_PROC_BIND=close, "
                   f"OMP_PLACES=cores), final relres {cpu['c1_relres']:.2e}; "
                   f"{'reference do_Hmvp compiled in place (gcc -O3 -fopenmp)' if cpu['use_ref'] else 'oracle stencil'}"
                   " driving a C restatement of FGMRES.m"),
        "host": cpu["host"],
        "omp": cpu["omp"],
        "c1": {"cpu_rhs_per_s": 1.0 / c1_s, "cpu_seconds": c1_s, "cpu_iters": cpu["c1_iters"],
               "gpu_rhs_per_s": 1.0 / gpu_s, "gpu_seconds": gpu_s, "gpu_iters": ig.iters,
               "gpu_over_cpu": c1_s / gpu_s},
        "c2_matvec": {"ms": cpu["c2_matvec_s"] * 1e3, "gpts": N2 / cpu["c2_matvec_s"] / 1e9,
                      "GBps": 40.0 * N2 / cpu["c2_matvec_s"] / 1e9},
        "c2_solve_estimate": {"rhs_per_s": 1.0 / (cpu["c2_sample_seconds"] / max(1, cpu["c2_sample_iters"]) *
                                                  gpu_c2_iters),
                              "sample": f"{cpu['c2_sample_iters']} Arnoldi steps of the C2 solve in "
                                        f"{cpu['c2_sample_seconds']:.1f} s, scaled by the GPU solve's "
                                        f"{gpu_c2_iters} iterations"},
    }


def mixed_leg(plain_rhs_per_s, restart=None):
    """8 sources on the benchmark model, batched mixed precision, versus the
    complex128 solutions of the same sources."""
    import torch
    from helm3d import BatchSolver, Helm3d
    from helm3d.models import config_c3, extend, wavenumber, vec, point_source
    cfg = config_c3()
    wn = vec(wavenumber(extend(cfg.v, cfg.npml), cfg.freq))
    dev = torch.device("cuda:0")
    H = Helm3d(cfg.nt, (cfg.h,) * 3, cfg.npml6, torch.from_numpy(wn).to(dev))
    bs = [torch.from_numpy(point_source(cfg, s)).to(dev) for s in cfg.sources]
    S = BatchSolver(dev)
    tuned = S.autotune([H] * len(bs), bs, inner_maxit=20)  # launch shape of the batched kernels
    kw = dict(method="mixed", tol=1e-6, maxinnerit=restart or MIXED_RESTART, inner_maxit=20)
    S.solve([H] * len(bs), bs, **kw)  # warm-up (allocation, graph capture)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xs, infos = S.solve([H] * len(bs), bs, **kw)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    errs = []
    for b, x in zip(bs, xs):
        xr, ir = H.solve(b, tol=1e-6)
        errs.append(float((x - xr).norm() / xr.norm()))
    rate = len(bs) / dt
    return {"workload": "8 point sources (z=5, y=100, x=round(linspace(20,179,8))) on the 240x240x140 benchmark "
                        "model, one batched mixed-precision solve to 1e-6 FP64 residual",
            "rhs_per_s": rate, "seconds": dt, "vs_plain": rate / plain_rhs_per_s if plain_rhs_per_s else None,
            "outer_iters": [i.iters for i in infos], "outer_restart": kw["maxinnerit"], "inner_iters_per_outer": 20,
            "final_relres": [i.relres for i in infos], "err_vs_complex128": errs, "launch": tuned,
            "dtype": "c64 inner / f64 outer"}


def uneven_leg(plain_rhs_per_s, restart=None):
    """The uneven batched workload: 8 subproblems (grids between 96^3 and
    256x256x128 + 20-pt PML, 3-8 Hz, 4-32 impulses each), every right-hand
    side solved to 1e-6 FP64 residual with batched mixed precision, split into
    batches that fit the device."""
    import torch
    from helm3d import BatchSolver, Helm3d
    from helm3d.models import config_c4, extend, wavenumber, vec
    subs = config_c4()
    dev = torch.device("cuda:0")
    ops, bs, owner = [], [], []
    for k, c in enumerate(subs):
        wn = vec(wavenumber(extend(c.v, c.npml), c.freq))
        H = Helm3d(c.nt, (c.h,) * 3, c.npml6, torch.from_numpy(wn).to(dev))
        p, (nx, ny, _) = c.npml, c.nt
        for s in c.sources:
            b = torch.zeros(H.N, dtype=torch.complex128, device=dev)
            b[(s[0] + p) + (s[1] + p) * nx + (s[2] + p) * nx * ny] = 1.0
            ops.append(H)
            bs.append(b)
            owner.append(k)
    S = BatchSolver(dev)
    first = [owner.index(k) for k in range(len(subs))]  # one right-hand side of every subproblem
    tuned = S.autotune([ops[i] for i in first], [bs[i] for i in first], inner_maxit=20, tile_rows=(4, 8),
                       prefetch=(2,), divisors=(1, 2))
    kw = dict(method="mixed", tol=1e-6, maxinnerit=restart or MIXED_RESTART, inner_maxit=20)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xs, infos, chunks = S.solve_chunked(ops, bs, **kw)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    rate = len(bs) / dt
    del xs
    torch.cuda.empty_cache()
    # the same right-hand sides with FGMRES + 4-level complex64 V-cycle, four
    # hierarchies per subproblem (set-up per subproblem not timed)
    from helm3d.mg import MGPool
    mg_t, mg_conv, mg_rel = 0.0, True, 0.0
    for k, c in enumerate(subs):
        pool = MGPool(vec(c.v), c.n_phys, c.h, c.npml, c.freq, nlevels=4, k=MG_K_TUNED, device=dev, workers=4)
        mine = [bs[i] for i in range(len(bs)) if owner[i] == k]
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        _, inf = pool.solve_many(mine, precision="mixed")
        torch.cuda.synchronize()
        mg_t += time.perf_counter() - t1
        mg_conv = mg_conv and all(i.converged for i in inf)
        mg_rel = max(mg_rel, max(i.relres for i in inf))
        pool.close()
        del pool
        torch.cuda.empty_cache()
    per_sub = []
    for k, c in enumerate(subs):
        its = [infos[i].iters for i in range(len(bs)) if owner[i] == k]
        per_sub.append({"n_phys": list(c.n_phys), "nt": list(c.nt), "freq": round(c.freq, 3), "rhs": len(its),
                        "outer_iters": [min(its), max(its)]})
    return {"workload": "8 subproblems, grids in [96^3, 256x256x128] + 20-pt PML, h=10 m, 3-8 Hz, 4-32 impulses each "
                        f"({len(bs)} right-hand sides), batched mixed-precision solves to 1e-6 FP64 residual",
            "rhs_per_s": rate, "seconds": dt,
            "vs_plain_c2": rate / plain_rhs_per_s if plain_rhs_per_s else None,
            "grid_point_rhs_per_s": sum(H.N for H in ops) / dt,
            "batches": [len(ch) for ch in chunks], "converged": all(i.converged for i in infos),
            "max_relres": max(i.relres for i in infos), "subproblems": per_sub, "launch": tuned,
            "dtype": "c64 inner / f64 outer",
            "multigrid": {"rhs_per_s": len(bs) / mg_t, "seconds": mg_t, "converged": mg_conv, "max_relres": mg_rel,
                          "solver": f"FGMRES(20) + 4-level complex64 ML_GMRES V-cycle, k = {MG_K_TUNED}, 4 streams "
                                    "per subproblem"}}


MG_K_DEFAULT = (3, 5, 3, 5)  # ML_GMRES.m:3-6
MIXED_RESTART = 8            # outer FGMRES restart of the batched mixed-precision legs (8, 10, 15, 20, 30 swept)
MG_K_TUNED = (1, 2, 3, 5)    # one smoothing cycle of 2 steps: fastest in sweeps of k on the 240x240x140 solve and the FWI gradient


def mg_leg(plain_rhs_per_s, solves=3, configs=((3, "c128", MG_K_DEFAULT), (3, "mixed", MG_K_DEFAULT),
                                               (4, "c128", MG_K_DEFAULT), (4, "mixed", MG_K_DEFAULT),
                                               (4, "mixed", MG_K_TUNED))):
    """The benchmark solve (240x240x140, 6 Hz, one impulse, 1e-6 FP64
    residual) with FGMRES(20) preconditioned by the ML_GMRES V-cycle: 3 levels
    (ML_GMRES.m) and 4 levels, the V-cycle in complex128 and in complex64,
    ML_GMRES's cycle sizes k and a tuned k."""
    import torch
    from helm3d.mg import MLGMRES
    from helm3d.models import config_c2, vec
    cfg = config_c2()
    dev = torch.device("cuda:0")
    runs = []
    for levels in sorted({cf[0] for cf in configs}):
        M = MLGMRES(vec(cfg.v), cfg.n_phys, cfg.h, cfg.npml, cfg.freq, nlevels=levels, device=dev)
        nx, ny, _ = M.grids[0]["nt"]
        p, s = cfg.npml, cfg.sources[0]
        b = torch.zeros(M.H.N, dtype=torch.complex128, device=dev)
        b[(s[0] + p) + (s[1] + p) * nx + (s[2] + p) * nx * ny] = 1.0
        for _, prec, k in [cf for cf in configs if cf[0] == levels]:
            M.k = tuple(k)
            M.solve(b, tol=1e-6, precision=prec)  # warm-up (workspaces)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(solves):
                x, info = M.solve(b, tol=1e-6, precision=prec)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / solves
            relres = float((b - M.H.apply(x)).norm() / b.norm())
            runs.append({"levels": levels, "grids": [list(g["nt"]) for g in M.grids], "k": list(k),
                         "vcycle": "complex128" if prec == "c128" else "complex64", "rhs_per_s": 1.0 / dt,
                         "seconds_per_solve": dt,
                         "vs_plain": (1.0 / dt) / plain_rhs_per_s if plain_rhs_per_s else None,
                         "outer_iters": info.iters, "final_relres": relres})
        M.close()
        del M
        torch.cuda.empty_cache()
    best = max(runs, key=lambda r: r["rhs_per_s"])
    return {"workload": "the 240x240x140 benchmark solve (one impulse, 6 Hz, 1e-6 FP64 residual) with FGMRES(20) "
                        "preconditioned by the ML_GMRES V-cycle",
            "rhs_per_s": best["rhs_per_s"], "best": {q: best[q] for q in ("levels", "vcycle", "k")}, "runs": runs,
            "dtype": "f64 outer / c128 or c64 V-cycle"}


def mg_sources_leg(plain_rhs_per_s, workers=4):
    """The 8 surface sources of the mixed-precision workload (C2 model) solved
    to 1e-6 FP64 residual by FGMRES with a 4-level complex64 V-cycle, four
    hierarchies on four streams; error of each solution against a complex128
    solve to 1e-10."""
    import torch
    from helm3d.mg import MGPool
    from helm3d.models import config_c3, vec
    cfg = config_c3()
    dev = torch.device("cuda:0")
    pool = MGPool(vec(cfg.v), cfg.n_phys, cfg.h, cfg.npml, cfg.freq, nlevels=4, k=MG_K_TUNED, device=dev,
                  workers=workers)
    nx, ny, _ = pool.grids[0]["nt"]
    p = cfg.npml
    bs = []
    for s in cfg.sources:
        b = torch.zeros(pool.H.N, dtype=torch.complex128, device=dev)
        b[(s[0] + p) + (s[1] + p) * nx + (s[2] + p) * nx * ny] = 1.0
        bs.append(b)
    pool.solve_many(bs, precision="mixed")  # warm-up: every hierarchy allocates its workspaces
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xs, infos = pool.solve_many(bs, precision="mixed")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    xr, _ = pool.solve_many(bs, tol=1e-10, precision="c128")
    errs = [float((x - r).norm() / r.norm()) for x, r in zip(xs, xr)]
    rel = [float((b - pool.H.apply(x)).norm() / b.norm()) for b, x in zip(bs, xs)]
    rate = len(bs) / dt
    pool.close()
    return {"workload": "8 point sources (z=5, y=100, x=round(linspace(20,179,8))) on the 240x240x140 benchmark "
                        "model, FGMRES(20) in FP64 + 4-level complex64 ML_GMRES V-cycle (inner FGMRES smoothers "
                        f"and coarse solves in complex64, k = {MG_K_TUNED}), {workers} hierarchies on {workers} "
                        "streams, 1e-6 FP64 residual",
            "rhs_per_s": rate, "seconds": dt, "vs_plain": rate / plain_rhs_per_s if plain_rhs_per_s else None,
            "outer_iters": [i.iters for i in infos], "final_relres": rel, "err_vs_complex128": errs,
            "dtype": "c64 V-cycle / f64 outer"}


def fwi_leg(solver="mg", workers=8):
    """One least-squares FWI gradient on the 256x256x128 workload; every
    solve to 1e-6 with FGMRES preconditioned by a 4-level complex64 V-cycle,
    eight solves at a time on eight streams (solver "mg"), or with batched
    mixed-precision solves ("mixed")."""
    import torch
    from helm3d.fwi import FWIOpts, FWIProblem
    from helm3d.models import fwi_c5
    model, m_true, m0 = fwi_c5()
    dev = torch.device("cuda:0")
    if solver == "mg":
        # 8 hierarchies: the 16 sources of a frequency split evenly
        opts = FWIOpts(pml=20, pml_max=20, tol=1e-6, solver="mg", mg_levels=4, mg_precision="mixed",
                       mg_workers=workers, mg_k=MG_K_TUNED)
        how = (f"FGMRES(20) + 4-level complex64 ML_GMRES V-cycle (k = {MG_K_TUNED}), FP64 outer, {workers} "
               f"solve{'s' if workers > 1 else ''} at a time on {workers} stream{'s' if workers > 1 else ''}")
    else:
        opts = FWIOpts(pml=20, pml_max=20, tol=1e-6, solver="mixed", batch=16, inner_iters=20)
        how = "batched mixed-precision solves (16 sources per batch)"
    P = FWIProblem(model, opts, device=dev)
    D = P.forw_model(m_true)  # observed data (not timed)
    P.opts.stats.clear()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f, g = P.objective(m0, D)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    its = [i.iters for i in P.opts.stats]
    return 60.0 / dt, {"workload": "LS-FWI gradient w.r.t. s2/m2, 256x256x128 + 20-pt PML at 12.5 m, 16 sources, "
                                   f"64x64 receivers, 4 and 6 Hz, solves to 1e-6: {how}",
                       "seconds_per_gradient": dt, "solves": len(its), "outer_iters_mean": sum(its) / len(its),
                       "misfit": f, "grad_norm": float(g.norm())}


def kacz_leg(sweeps=5):
    """The Kaczmarz path (Helm3dkaczswp_mex.c, CGMN.m, CRMN.m): double sweeps
    of the 240x240x140 benchmark model timed with events (forward and adjoint
    rows), and CGMN and CRMN solves of the C1 problem (64^3 + 10-pt PML, 5 Hz)
    to 1e-6.  Algorithmic bytes of a double sweep: each of the 54 colour
    launches reads and writes x at every point and reads wn (40 B/pt), and
    each row reads b once per direction (2 x 16 B/pt): 2192 B/pt."""
    import numpy as np
    import torch
    from helm3d import Helm3d
    from helm3d.models import config_c1, config_c2, extend, point_source, vec, wavenumber
    dev = torch.device("cuda:0")
    cfg = config_c2()
    wn = vec(wavenumber(extend(cfg.v, cfg.npml), cfg.freq))
    H = Helm3d(cfg.nt, (cfg.h,) * 3, cfg.npml6, torch.from_numpy(wn).to(dev))
    rng = np.random.default_rng(3)
    x = torch.from_numpy(rng.standard_normal(H.N) + 1j * rng.standard_normal(H.N)).to(dev)
    b = torch.from_numpy(rng.standard_normal(H.N) + 1j * rng.standard_normal(H.N)).to(dev)
    alg = 2192.0 * H.N
    out = {"workload": "double Kaczmarz sweeps (27-colour order, omega 1.5) on the 240x240x140 benchmark model; "
                       "CGMN and CRMN solves of C1 (64^3 + 10-pt PML, 5 Hz, one impulse) to 1e-6",
           "alg_bytes_per_double_sweep": alg}
    for mode, name in ((1, "forward"), (2, "adjoint")):
        H.kacz_sweep(x, b, mode=mode)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        H.kacz_sweep(x, b, mode=mode, nsweeps=sweeps)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / sweeps
        out[name] = {"ms_per_double_sweep": ms, "GB/s": alg / (ms * 1e-3) / 1e9,
                     "frac_of_hbm_peak": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    del H, x, b
    torch.cuda.empty_cache()
    c1 = config_c1()
    wn1 = vec(wavenumber(extend(c1.v, c1.npml), c1.freq))
    H1 = Helm3d(c1.nt, (c1.h,) * 3, c1.npml6, torch.from_numpy(wn1).to(dev))
    b1 = torch.from_numpy(point_source(c1, c1.sources[0])).to(dev)
    for method in ("cgmn", "crmn"):
        H1.solve_cgmn(b1, tol=1e-6, maxit=20000, method=method)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, info = H1.solve_cgmn(b1, tol=1e-6, maxit=20000, method=method)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[method + "_c1"] = {"rhs_per_s": 1.0 / dt, "seconds": dt, "iters": info.iters,
                               "relres": info.relres, "converged": info.converged}
    return out


def copy_mode(args):
    """K copies of a 256 MiB buffer: the known-size kernel that calibrates the
    FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    nbytes = 256 << 20
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(args.warmup):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    print(json.dumps({"mode": "copy", "bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
                      "ms_per_launch": dt * 1e3, "GB/s": 2 * nbytes / dt / 1e9}))


def copy_ceiling_gbs(nbytes=1 << 30, reps=10):
    """Measured device-copy bandwidth (read + write bytes / time) of a 1 GiB
    buffer, timed with events on the current stream: the practical ceiling
    beside the 8 TB/s peak."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * nbytes / (ms * 1e-3) / 1e9


def matvec_mode(args, cfg, wn):
    """K FP64 forward matvecs y = H x on the benchmark grid (40 B/pt)."""
    import numpy as np
    import torch
    from helm3d import Helm3d
    dev = torch.device("cuda:0")
    H = Helm3d(cfg.nt, (cfg.h,) * 3, cfg.npml6, torch.from_numpy(wn).to(dev))
    if args.launch:
        H.set_launch(*(int(a) for a in args.launch.split(",")))
    elif not args.no_tune:
        H.autotune(mode=1, precision="c128")
    x = torch.from_numpy(np.random.default_rng(0).standard_normal(H.N) + 0j).to(dev)
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        H.apply(x, out=y)
    torch.cuda.synchronize()
    H.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        H.apply(x, out=y)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    ms, n = H.profile_read()["stencil"]
    H.profile(False)
    kern = ms / max(1, n)
    gbs = 40.0 * H.N / (kern * 1e-3) / 1e9
    print(json.dumps({"mode": "matvec", "grid_points": H.N, "alg_bytes_per_launch": 40.0 * H.N,
                      "kernel_ms": kern, "wall_ms_per_launch": dt * 1e3, "GB/s": gbs,
                      "pct_hbm_peak": 100.0 * gbs / HBM_PEAK_GBS, "launch": H.get_launch()}))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the batched mixed-precision leg")
    ap.add_argument("--no-fwi", action="store_true", help="skip the FWI gradient leg")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven batched leg")
    ap.add_argument("--no-mg", action="store_true", help="skip the multigrid-preconditioned leg")
    ap.add_argument("--no-kacz", action="store_true", help="skip the Kaczmarz / CGMN leg")
    ap.add_argument("--no-tune", action="store_true", help="keep the default stencil launch shape")
    ap.add_argument("--launch", default=None,
                    help="--mode matvec only: stencil launch shape 'kchunk,tile_rows,prefetch[,rows_per_thread]' "
                         "(no autotuning)")
    ap.add_argument("--cpu-sample-iters", type=int, default=20)
    ap.add_argument("--fwi-workers", type=int, default=8,
                    help="--mode fwi only: multigrid hierarchies solving at once (host threads)")
    ap.add_argument("--mixed-restart", type=int, default=None,
                    help="outer FGMRES restart of the batched mixed-precision legs (default MIXED_RESTART)")
    ap.add_argument("--mg-config", default=None,
                    help="--mode mg only: one 'levels,precision[,k1,k2,k3,k4]' set (e.g. 4,mixed or "
                         "4,mixed,1,5,3,5) instead of the default list")
    ap.add_argument("--mode", choices=("solve", "matvec", "copy", "mixed", "uneven", "fwi", "mg", "mg-sources",
                                       "kacz", "cpu-baseline"),
                    default="solve",
                    help="solve: the benchmark; matvec: K FP64 forward stencil launches at the same grid "
                         "(profiling workload); copy: K device copies of 256 MiB (PMC calibration); "
                         "mixed / uneven / fwi / mg / mg-sources / kacz: one of the side legs alone (profiling "
                         "workloads); "
                         "cpu-baseline: the CPU part of the baseline alone (a child process of the default run)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: the hot path is one solve per GPU and this build runs on a single "
              f"MI355X (got --gpus {args.gpus}, WORLD_SIZE {world}); multi-GPU sharding over (source, frequency) "
              f"is not built.", file=sys.stderr)
        sys.exit(2)

    if args.mode == "cpu-baseline":
        return cpu_mode(args)  # no torch in this process (see _cpu_env)
    sys.path.insert(0, PKG)
    import numpy as np
    import torch
    from helm3d import Helm3d
    from helm3d.models import config_c2, extend, wavenumber, vec, point_source

    if args.mode == "copy":
        return copy_mode(args)
    if args.mode == "kacz":
        print(json.dumps({"mode": "kacz", **kacz_leg()}))
        return
    if args.mode in ("mixed", "uneven", "fwi", "mg", "mg-sources"):
        plain_rhs_per_s = None  # no plain solve in this run: ratios are left out
        if args.mode == "mixed":
            res = mixed_leg(plain_rhs_per_s, args.mixed_restart)
        elif args.mode == "uneven":
            res = uneven_leg(plain_rhs_per_s, args.mixed_restart)
        elif args.mode == "mg-sources":
            res = mg_sources_leg(plain_rhs_per_s)
        elif args.mode == "mg":
            if args.mg_config:
                f = args.mg_config.split(",")
                k = tuple(int(a) for a in f[2:6]) if len(f) >= 6 else MG_K_DEFAULT
                res = mg_leg(plain_rhs_per_s, solves=args.steps, configs=((int(f[0]), f[1], k),))
            else:
                res = mg_leg(plain_rhs_per_s)
        else:
            res = dict(zip(("grads_per_min", "fwi"), fwi_leg(workers=args.fwi_workers)))
        print(json.dumps({"mode": args.mode, **res}))
        return
    cfg = config_c2()
    wn = vec(wavenumber(extend(cfg.v, cfg.npml), cfg.freq))
    if args.mode == "matvec":
        return matvec_mode(args, cfg, wn)
    b = point_source(cfg, cfg.sources[0])
    dev = torch.device("cuda:0")
    H = Helm3d(cfg.nt, (cfg.h,) * 3, cfg.npml6, torch.from_numpy(wn).to(dev))
    bt = torch.from_numpy(b).to(dev)
    N = H.N
    m = 20
    tuned = H.autotune(mode=1, precision="c128") if not args.no_tune else H.get_launch()

    for _ in range(args.warmup):
        x, info = H.solve(bt, tol=1e-6, maxinnerit=m)
        assert info.converged, info
    torch.cuda.synchronize()

    # timed region (value): K solves, nothing instrumented
    infos = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        x, info = H.solve(bt, tol=1e-6, maxinnerit=m)
        infos.append(info)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    for info in infos:
        assert info.converged and info.relres < 1e-6, info
    # the same K solves again with every launch bracketed by HIP events on
    # the library's stream: kernel breakdown and rooflines
    H.profile(True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        xp, ip = H.solve(bt, tol=1e-6, maxinnerit=m)
        assert ip.converged and ip.iters == infos[-1].iters, ip
    torch.cuda.synchronize()
    elapsed_prof = time.perf_counter() - t1
    prof = H.profile_read()
    H.profile(False)

    iters = infos[-1].iters
    cycles = infos[-1].cycles
    st_ms, st_n = prof["stencil"]
    per_solve_launches = st_n / args.steps
    breakdown = {}
    for kind, (ms, n) in prof.items():
        byt = _alg_bytes(kind, iters, cycles, m, N, per_solve_launches) * args.steps
        breakdown[kind] = {"ms_per_solve": ms / args.steps, "launches_per_solve": n / args.steps,
                           "alg_bytes": byt, "total_ms": ms, "launches": n,
                           "GB/s": (byt / (ms * 1e-3) / 1e9) if ms > 0 and byt > 0 else None}

    def roof(kind, kernel, profile):
        b = breakdown[kind]
        ach = b["alg_bytes"] / (b["total_ms"] * 1e-3) / 1e9
        nl = max(1, b["launches"])
        tr = _traffic_from_profile(profile)
        return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": tr, "kernel": kernel,
                "alg_bytes_per_launch": b["alg_bytes"] / nl, "avg_launch_ms": b["total_ms"] / nl,
                "share_of_solve_kernel_time": b["total_ms"] / sum(v["total_ms"] for v in breakdown.values())}

    dominant = max(("update", "dots", "stencil"), key=lambda k: breakdown[k]["total_ms"])
    kernels = {"update": "update_kernel<double,K>: CGS update w -= V h and ||w||^2, 16 (K+2) B/pt",
               "dots": "dots_kernel<double,K>: CGS dots V^H w, 16 (K+1) B/pt",
               "stencil": "helm27 stencil, FP64 forward matvec (40 B/pt) + fused residual (56 B/pt)"}
    profiles = {"update": "gs_update_c2_pmc.txt", "dots": "gs_dots_c2_pmc.txt", "stencil": "stencil_fwd_c2_pmc.txt"}
    roofline = roof(dominant, kernels[dominant], profiles[dominant])
    stencil_roofline = roof("stencil", kernels["stencil"], "stencil_fwd_c2_pmc.txt")
    copy_gbs = copy_ceiling_gbs()
    for r in (roofline, stencil_roofline):
        r["measured_copy_GBps"] = copy_gbs
        r["frac_of_measured_copy"] = r["achieved"] / copy_gbs
    stencil_roofline["note"] = ("near the FP64 vector ridge (78.6 TFLOP/s / 8 TB/s = 10 flop/B): about 176 VALU "
                                "instructions (mostly FP64) and 16 LDS instructions per point against 40 B; the "
                                "arithmetic does not fully overlap memory (a copy of the kernel without it runs at "
                                "4.8-5.3 TB/s), and more planes of loads in flight cost registers")
    for v in breakdown.values():
        for k in ("alg_bytes", "total_ms", "launches"):
            v.pop(k)

    out = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "plain complex128 FGMRES(20) solve to 1e-6 of the 27-pt Helmholtz stencil with PML, "
                        "200x200x100 + 20-pt PML (240x240x140), h=10 m, 6 Hz, 1 point source",
            "grid_points": N, "restart": m, "tol": 1e-6, "precision": "complex128", "stencil_launch": tuned,
        },
        "roofline": roofline,
        "stencil_roofline": stencil_roofline,
        "matvec_pct_hbm_peak": 100.0 * stencil_roofline["frac"],
        "iters_per_solve": iters,
        "cycles_per_solve": cycles,
        "final_relres": infos[-1].relres,
        "kernel_breakdown": breakdown,
        "instrumented_rhs_per_s": args.steps / elapsed_prof,
        "fwi_grads_per_min": None,
    }
    del H
    torch.cuda.empty_cache()
    if not args.no_mixed:
        out["mixed_batched"] = mixed_leg(args.steps / elapsed, args.mixed_restart)
        print(f"[bench] mixed leg done: {out['mixed_batched']['rhs_per_s']:.3f} RHS/s", file=sys.stderr, flush=True)
    if not args.no_mg:
        out["multigrid"] = mg_leg(args.steps / elapsed)
        print(f"[bench] multigrid leg done: {out['multigrid']['rhs_per_s']:.3f} RHS/s", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
        out["multigrid_sources"] = mg_sources_leg(args.steps / elapsed)
        print(f"[bench] multigrid sources leg done: {out['multigrid_sources']['rhs_per_s']:.3f} RHS/s",
              file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    if not args.no_uneven:
        out["batched_uneven"] = uneven_leg(args.steps / elapsed, args.mixed_restart)
        print(f"[bench] uneven batched leg done: {out['batched_uneven']['rhs_per_s']:.3f} RHS/s", file=sys.stderr,
              flush=True)
        torch.cuda.empty_cache()
    if not args.no_kacz:
        out["kaczmarz"] = kacz_leg()
        print(f"[bench] kaczmarz leg done: {out['kaczmarz']['forward']['ms_per_double_sweep']:.2f} ms per double "
              f"sweep", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    if not args.no_fwi:
        out["fwi_grads_per_min"], out["fwi"] = fwi_leg()
        print(f"[bench] fwi leg done: {out['fwi_grads_per_min']:.3f} grads/min", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(iters, args.cpu_sample_iters)
        print(f"[bench] cpu baseline done: C1 {out['cpu_baseline']['c1']['cpu_seconds']:.1f} s on "
              f"{out['cpu_baseline']['cores']} threads", file=sys.stderr, flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
