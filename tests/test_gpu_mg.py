"""GPU multigrid-preconditioned FGMRES (ML_GMRES) against the CPU oracle.

Bars: the device transfers equal the oracle's trilinear prolongation and 1/8
of its transpose to 1e-13; the preconditioned solve, with the V-cycle in
complex128 or in complex64 (mixed), reaches the tolerance with the true
residual checked by the oracle operator in FP64, in far fewer outer
iterations than plain FGMRES(20), and its solution matches the oracle's
FGMRES solution (forward and adjoint)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _model():
    from helm3d.models import config_small
    cfg = config_small(n=(40, 36, 30), npml=8, freq=6.0)
    return cfg


def test_transfers_match_oracle():
    import torch
    from helm3d.mg import MLGMRES
    from helm3d.models import vec
    cfg = _model()
    M = MLGMRES(vec(cfg.v), cfg.n_phys, cfg.h, cfg.npml, cfg.freq, nlevels=3)
    rng = np.random.default_rng(0)
    for lvl in range(2):
        nf, nc = M.grids[lvl]["nt"], M.grids[lvl + 1]["nt"]
        P = O.prolongation(nf, nc)
        xc = rng.standard_normal(P.shape[1]) + 1j * rng.standard_normal(P.shape[1])
        xf = rng.standard_normal(P.shape[0]) + 1j * rng.standard_normal(P.shape[0])
        yf = M.transfer(lvl, torch.from_numpy(xc).cuda(), "prolong").cpu().numpy()
        yc = M.transfer(lvl, torch.from_numpy(xf).cuda(), "restrict").cpu().numpy()
        assert np.linalg.norm(yf - P @ xc) <= 1e-13 * np.linalg.norm(P @ xc)
        assert np.linalg.norm(yc - (P.T @ xf) / 8) <= 1e-13 * np.linalg.norm(P.T @ xf / 8)


@pytest.mark.parametrize("precision", ["c128", "mixed"])
@pytest.mark.parametrize("mode", [1, 2])
def test_mg_solve(mode, precision):
    import torch
    from helm3d.mg import MLGMRES
    from helm3d.models import extend, point_source, vec, wavenumber
    cfg = _model()
    M = MLGMRES(vec(cfg.v), cfg.n_phys, cfg.h, cfg.npml, cfg.freq, nlevels=3)
    print("levels", [g["nt"] for g in M.grids], flush=True)
    b = point_source(cfg, cfg.sources[0])
    x, info = M.solve(torch.from_numpy(b).cuda(), mode=mode, tol=1e-6, maxinnerit=20, precision=precision)
    wn = vec(wavenumber(extend(cfg.v, cfg.npml), cfg.freq))
    v = "forw" if mode == 1 else "adj"
    h3, p6 = (cfg.h,) * 3, np.array(cfg.npml6)
    xn = x.cpu().numpy()
    r = b - O.hmvp(v, wn, h3, cfg.nt, p6, xn)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    xo, io = O.fgmres(v, wn, h3, cfg.nt, p6, b, tol=1e-8, maxinnerit=20)
    err = np.linalg.norm(xn - xo) / np.linalg.norm(xo)
    print(f"mg mode={mode} {precision}: iters={info.iters} cycles={info.cycles} relres={info.relres:.2e} (oracle {rel:.2e}) "
          f"plain FGMRES iters={io['iters']} err={err:.2e}", flush=True)
    assert info.converged and rel < 1e-6 and abs(rel - info.relres) < 1e-9
    assert info.iters * 5 < io["iters"]
    assert err < 1e-4


@pytest.mark.parametrize("k", [(1, 2, 3, 5), (2, 8, 2, 7), (1, 12, 3, 3)])
def test_mg_solve_cycle_sizes(k):
    """Smoother and coarse-solve cycle sizes k = (ks_outer, ks_inner,
    kc_outer, kc_inner) of ML_GMRES.m: the smoothers and coarsest solves run
    their restarted GMRES decisions on the device, with the least squares in
    registers (cycles of up to 5 steps) or in LDS (longer cycles); every
    setting reaches the tolerance with the true residual checked by the
    oracle operator."""
    import torch
    from helm3d.mg import MLGMRES
    from helm3d.models import extend, point_source, vec, wavenumber
    cfg = _model()
    M = MLGMRES(vec(cfg.v), cfg.n_phys, cfg.h, cfg.npml, cfg.freq, nlevels=3, k=k)
    b = point_source(cfg, cfg.sources[0])
    x, info = M.solve(torch.from_numpy(b).cuda(), tol=1e-6, maxinnerit=20, precision="mixed")
    wn = vec(wavenumber(extend(cfg.v, cfg.npml), cfg.freq))
    xn = x.cpu().numpy()
    r = b - O.hmvp("forw", wn, (cfg.h,) * 3, cfg.nt, np.array(cfg.npml6), xn)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    print(f"k={k}: iters={info.iters} relres={info.relres:.2e} (oracle {rel:.2e})", flush=True)
    assert info.converged and rel < 1e-6 and abs(rel - info.relres) < 1e-9
    M.close()


# (n_phys, npml, levels, free surface, k) and the grids each hierarchy builds;
# every coarsened level keeps the 4 physical points per axis that the cubic
# model coarsening needs
TINY = [
    ((7, 8, 7), 1, 3, False, (1, 2, 3, 5)),       # 9x10x9, 6x6x6, 4x4x4
    ((13, 13, 13), 2, 4, True, (3, 5, 3, 5)),     # 17x17x15, 9x9x8, 6x6x5, 4x4x3
    ((4, 5, 4), 1, 2, False, (1, 30, 2, 30)),     # 6x7x6, 4x5x4: 30-step cycles on 80 to 252 points
]


@pytest.mark.parametrize("case", range(len(TINY)))
def test_tiny_hierarchies(case):
    """The smallest hierarchies the coarsening allows, down to 3 or 4 points
    per axis with the PML: the transfers against the oracle's P, and the
    V-cycle solve in both precisions and directions (device-driven cycles of
    up to 30 steps on vectors of 48 to 4335 points, shorter than one block of
    threads on the coarsest levels) reaching the tolerance with the oracle's
    operator and the oracle's solution (condition numbers 10 to 84, so the
    1e-6 residual bounds the error by about 1e-4)."""
    import torch
    from helm3d.mg import MLGMRES
    n, npml, nl, fs, k = TINY[case]
    rng = np.random.default_rng(40 + case)
    v = 2000.0 + 200.0 * rng.standard_normal(int(np.prod(n)))
    M = MLGMRES(v, n, 10.0, npml, 6.0, nlevels=nl, k=k, free_surface=fs)
    print(f"tiny hierarchy {[g['nt'] for g in M.grids]} k={k}", flush=True)
    for lvl in range(nl - 1):
        nf, nc = M.grids[lvl]["nt"], M.grids[lvl + 1]["nt"]
        P = O.prolongation(nf, nc)
        xc = rng.standard_normal(P.shape[1]) + 1j * rng.standard_normal(P.shape[1])
        xf = rng.standard_normal(P.shape[0]) + 1j * rng.standard_normal(P.shape[0])
        yf = M.transfer(lvl, torch.from_numpy(xc).cuda(), "prolong").cpu().numpy()
        yc = M.transfer(lvl, torch.from_numpy(xf).cuda(), "restrict").cpu().numpy()
        assert np.linalg.norm(yf - P @ xc) <= 1e-13 * np.linalg.norm(P @ xc)
        assert np.linalg.norm(yc - (P.T @ xf) / 8) <= 1e-13 * np.linalg.norm(P.T @ xf / 8)
    wn = M.H.wn.cpu().numpy()
    nt, p6, h3 = M.grids[0]["nt"], np.array(M.grids[0]["p6"]), (10.0,) * 3
    N = int(np.prod(nt))
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    for mode in (1, 2):
        variant = "forw" if mode == 1 else "adj"
        xo, _ = O.fgmres(variant, wn, h3, nt, p6, b, tol=1e-10, maxinnerit=20)
        for precision in ("c128", "mixed"):
            x, info = M.solve(torch.from_numpy(b).cuda(), mode=mode, tol=1e-6, maxit=200, precision=precision)
            xn = x.cpu().numpy()
            assert np.isfinite(xn).all()
            rel = np.linalg.norm(b - O.hmvp(variant, wn, h3, nt, p6, xn)) / np.linalg.norm(b)
            err = np.linalg.norm(xn - xo) / np.linalg.norm(xo)
            print(f"  mode={mode} {precision}: iters={info.iters} relres={info.relres:.2e} (oracle {rel:.2e}) "
                  f"err={err:.2e}", flush=True)
            assert info.converged and rel < 1e-6 and abs(rel - info.relres) < 1e-9
            assert err < 5e-4
    M.close()


def test_pool_of_hierarchies_matches_single():
    """Right-hand sides solved three at a time on separate streams (MGPool)
    give the solutions of one-at-a-time solves (to rounding: the hierarchies'
    launch shapes are tuned separately, which can reorder the block sums of a
    restart's residual norm)."""
    import torch
    from helm3d.mg import MGPool, MLGMRES
    from helm3d.models import point_source, vec
    cfg = _model()
    bs = [torch.from_numpy(point_source(cfg, s)).cuda() for s in [(5, 5, 3), (20, 18, 3), (30, 10, 20), (12, 30, 8),
                                                                  (33, 4, 3)]]
    M = MLGMRES(vec(cfg.v), cfg.n_phys, cfg.h, cfg.npml, cfg.freq, nlevels=3)
    ref = [M.solve(b, tol=1e-6, precision="mixed") for b in bs]
    pool = MGPool(vec(cfg.v), cfg.n_phys, cfg.h, cfg.npml, cfg.freq, nlevels=3, workers=3)
    xs, infos = pool.solve_many(bs, tol=1e-6, precision="mixed")
    print("pool iterations", [i.iters for i in infos], flush=True)
    for (xr, ir), x, inf in zip(ref, xs, infos):
        assert inf.converged and inf.iters == ir.iters
        assert float((x - xr).norm() / xr.norm()) <= 1e-12
