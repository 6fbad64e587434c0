"""GPU batched solves (helm3d_batch_solve) against the CPU oracle.

Bars: every item reaches ||b - A x|| / ||b|| < tol measured with the oracle
operator in FP64; the plain method (FGMRES_MT.m semantics) keeps all items in
lockstep (equal iteration counts) and, for one item, takes the iteration count
of the oracle's FGMRES.m restatement (within 2% + 2); mixed-precision items
stay within 1e-3 of the complex128 solution; a zero right-hand side returns
zeros; repeated batch solves are bit-identical.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _op(n, npml, freq, seed):
    import torch
    from helm3d import Helm3d
    from helm3d.models import config_small, extend, wavenumber, vec
    cfg = config_small(n=n, npml=npml, freq=freq, seed=seed)
    wn = vec(wavenumber(extend(cfg.v, cfg.npml), cfg.freq))
    H = Helm3d(cfg.nt, (cfg.h,) * 3, cfg.npml6, torch.from_numpy(wn).to("cuda:0"))
    return cfg, wn, H


def _rhs(cfg, src):
    from helm3d.models import point_source
    return point_source(cfg, src)


def _relres(cfg, wn, b, x, mode):
    v = "forw" if mode == 1 else "adj"
    r = b - O.hmvp(v, wn, (cfg.h,) * 3, cfg.nt, np.array(cfg.npml6), x)
    return np.linalg.norm(r) / np.linalg.norm(b)


def test_plain_single_item_matches_fgmres():
    import torch
    from helm3d import BatchSolver
    cfg, wn, H = _op((32, 32, 24), 8, 5.0, 11)
    b = _rhs(cfg, cfg.sources[0])
    print("plain single", cfg.nt, flush=True)
    _, info_o = O.fgmres("forw", wn, (cfg.h,) * 3, cfg.nt, np.array(cfg.npml6), b, tol=1e-6, maxinnerit=20)
    S = BatchSolver("cuda:0")
    xs, infos = S.solve([H], [torch.from_numpy(b).cuda()], method="plain", tol=1e-6)
    x = xs[0].cpu().numpy()
    rel = _relres(cfg, wn, b, x, 1)
    print("batch iters", infos[0].iters, "oracle", info_o["iters"], "relres", rel, infos[0].relres)
    assert infos[0].converged and rel < 1e-6
    assert abs(rel - infos[0].relres) <= 1e-9
    assert abs(infos[0].iters - info_o["iters"]) <= 0.02 * info_o["iters"] + 2


@pytest.mark.parametrize("mode", [1, 2])
def test_plain_lockstep_uneven_grids(mode):
    import torch
    from helm3d import BatchSolver
    c1, w1, H1 = _op((32, 32, 24), 8, 5.0, 12)
    c2, w2, H2 = _op((40, 28, 30), 6, 6.0, 13)
    items = [(c1, w1, H1, (16, 16, 3)), (c1, w1, H1, (5, 20, 10)), (c2, w2, H2, (20, 14, 3))]
    bs = [_rhs(c, s) for c, _, _, s in items]
    print("plain lockstep mode", mode, flush=True)
    S = BatchSolver("cuda:0")
    xs, infos = S.solve([it[2] for it in items], [torch.from_numpy(b).cuda() for b in bs], mode=mode,
                        method="plain", tol=1e-6)
    for (c, w, _, _), b, x, inf in zip(items, bs, xs, infos):
        rel = _relres(c, w, b, x.cpu().numpy(), mode)
        print("item", inf.iters, inf.cycles, inf.relres, rel)
        assert inf.converged and rel < 1e-6
    assert len({inf.iters for inf in infos}) == 1  # lockstep (FGMRES_MT.m)


@pytest.mark.parametrize("method", ["plain", "mixed"])
def test_batch_with_degenerate_item(method):
    """A batch holding a grid of 3 x 1 x 2 points, whose Krylov space runs out
    before the tolerance (a breakdown ends its cycles), beside a 48 x 48 x 40
    grid: both items converge and the small one stays finite."""
    import torch
    from helm3d import BatchSolver, Helm3d
    c1, w1, H1 = _op((32, 32, 24), 8, 5.0, 12)
    b1 = _rhs(c1, (16, 16, 3))
    n2, p2, h2 = (3, 1, 2), np.array([[1, 1, 1], [1, 1, 1]]), (10.0, 10.0, 10.0)
    rng = np.random.default_rng(21)
    w2 = rng.uniform(2e-4, 2e-3, 6)
    b2 = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    H2 = Helm3d(n2, h2, p2, torch.from_numpy(w2).to("cuda:0"))
    print("batch with a degenerate item", method, flush=True)
    S = BatchSolver("cuda:0")
    extra = {"inner_maxit": 10} if method == "mixed" else {}
    xs, infos = S.solve([H1, H2], [torch.from_numpy(b1).cuda(), torch.from_numpy(b2).cuda()], mode=1,
                        method=method, tol=1e-6, **extra)
    r1 = _relres(c1, w1, b1, xs[0].cpu().numpy(), 1)
    x2 = xs[1].cpu().numpy()
    assert np.isfinite(x2).all()
    r2 = np.linalg.norm(b2 - O.hmvp("forw", w2, h2, n2, p2, x2)) / np.linalg.norm(b2)
    print("items", [(i.iters, i.cycles, i.relres) for i in infos], r1, r2, flush=True)
    assert infos[0].converged and infos[1].converged
    assert r1 < 1e-6 and r2 < 1e-6


@pytest.mark.parametrize("mode", [1, 2])
def test_mixed_batch_uneven_grids(mode):
    import torch
    from helm3d import BatchSolver
    c1, w1, H1 = _op((32, 32, 24), 8, 5.0, 14)
    c2, w2, H2 = _op((48, 36, 28), 8, 7.0, 15)
    items = [(c1, w1, H1, (16, 16, 3)), (c2, w2, H2, (24, 18, 3)), (c2, w2, H2, (8, 30, 20)),
             (c1, w1, H1, (3, 3, 3))]
    bs = [_rhs(c, s) for c, _, _, s in items]
    print("mixed batch mode", mode, flush=True)
    S = BatchSolver("cuda:0")
    xs, infos = S.solve([it[2] for it in items], [torch.from_numpy(b).cuda() for b in bs], mode=mode,
                        method="mixed", tol=1e-6, inner_maxit=10)
    for (c, w, H, _), b, x, inf in zip(items, bs, xs, infos):
        xn = x.cpu().numpy()
        rel = _relres(c, w, b, xn, mode)
        xr, _ = H.solve(torch.from_numpy(b).cuda(), mode=mode, tol=1e-10)
        err = np.linalg.norm(xn - xr.cpu().numpy()) / np.linalg.norm(xr.cpu().numpy())
        print("item iters", inf.iters, "inner", inf.inner_iters, "relres", inf.relres, rel, "err", err)
        assert inf.converged and rel < 1e-6
        assert err < 1e-3


@pytest.mark.parametrize("method", ["mixed", "plain"])
@pytest.mark.parametrize("mode", [1, 2])
def test_batch_two_rows_per_thread(mode, method):
    """Batched solves on the 64 x 8, two-rows-per-thread launch shape (the
    fused BiCGStab passes and the FP64 work-list stencils with two output rows
    per thread): every item reaches tol with the oracle operator and agrees
    with a complex128 solve, on grids whose y extent is not a multiple of 8."""
    import torch
    from helm3d import BatchSolver
    c1, w1, H1 = _op((32, 30, 24), 8, 5.0, 14)
    c2, w2, H2 = _op((48, 37, 28), 8, 7.0, 15)
    H1.set_launch(0, 8, 2, 2)
    H2.set_launch(13, 8, 2, 2)
    items = [(c1, w1, H1, (16, 16, 3)), (c2, w2, H2, (24, 18, 3)), (c2, w2, H2, (8, 30, 20))]
    bs = [_rhs(c, s) for c, _, _, s in items]
    print("batch two rows per thread, mode", mode, method, flush=True)
    S = BatchSolver("cuda:0")
    xs, infos = S.solve([it[2] for it in items], [torch.from_numpy(b).cuda() for b in bs], mode=mode,
                        method=method, tol=1e-6, inner_maxit=10)
    for (c, w, H, _), b, x, inf in zip(items, bs, xs, infos):
        xn = x.cpu().numpy()
        rel = _relres(c, w, b, xn, mode)
        xr, _ = H.solve(torch.from_numpy(b).cuda(), mode=mode, tol=1e-10)
        err = np.linalg.norm(xn - xr.cpu().numpy()) / np.linalg.norm(xr.cpu().numpy())
        print("item iters", inf.iters, "relres", inf.relres, rel, "err", err, flush=True)
        assert inf.converged and rel < 1e-6
        assert err < 1e-3


@pytest.mark.parametrize("mode", [1, 2])
def test_mixed_batch_complex_wavenumber(mode):
    """Fused complex64 BiCGStab on the complex-wn operators (WNC variants):
    attenuating wn = (1 + 0.05i) wn, residual checked with the oracle's wnc
    stencil."""
    import torch
    from helm3d import BatchSolver, Helm3d
    from helm3d.models import config_small, extend, wavenumber, vec
    cfg = config_small(n=(32, 30, 24), npml=8, freq=5.0, seed=18)
    wn = vec(wavenumber(extend(cfg.v, cfg.npml), cfg.freq)) * (1.0 + 0.05j)
    H = Helm3d(cfg.nt, (cfg.h,) * 3, cfg.npml6, torch.from_numpy(wn).to("cuda:0"))
    bs = [_rhs(cfg, s) for s in ((16, 15, 3), (5, 20, 12))]
    print("mixed batch wnc mode", mode, flush=True)
    S = BatchSolver("cuda:0")
    xs, infos = S.solve([H, H], [torch.from_numpy(b).cuda() for b in bs], mode=mode, method="mixed", tol=1e-6,
                        inner_maxit=10)
    v = "forw_wnc" if mode == 1 else "adj_wnc"
    for b, x, inf in zip(bs, xs, infos):
        xn = x.cpu().numpy()
        r = b - O.hmvp(v, wn, (cfg.h,) * 3, cfg.nt, np.array(cfg.npml6), xn)
        rel = np.linalg.norm(r) / np.linalg.norm(b)
        xr, _ = H.solve(torch.from_numpy(b).cuda(), mode=mode, tol=1e-10)
        err = np.linalg.norm(xn - xr.cpu().numpy()) / np.linalg.norm(xr.cpu().numpy())
        print("item iters", inf.iters, "relres", inf.relres, rel, "err", err)
        assert inf.converged and rel < 1e-6
        assert err < 1e-3


def test_mixed_single_matches_unbatched_mixed():
    """A one-item batch (fused complex64 BiCGStab) and the single-vector mixed
    solve (unfused kernels) converge to the same solution."""
    import torch
    from helm3d import BatchSolver
    cfg, wn, H = _op((32, 32, 24), 8, 5.0, 16)
    b = torch.from_numpy(_rhs(cfg, cfg.sources[0])).cuda()
    S = BatchSolver("cuda:0")
    xs, infos = S.solve([H], [b], method="mixed", tol=1e-8, inner_maxit=12)
    x1, i1 = H.solve_mixed(b, tol=1e-8, inner_tol=0.0, inner_maxit=12)
    print("batched", infos[0].iters, "single", i1.iters, flush=True)
    d = (xs[0] - x1).norm() / x1.norm()
    assert infos[0].converged and i1.converged
    assert float(d) < 1e-6
    # the fused inner solve is the same iteration: same outer step count
    assert abs(infos[0].iters - i1.iters) <= 1


def test_zero_rhs_and_determinism():
    import torch
    from helm3d import BatchSolver
    cfg, wn, H = _op((32, 32, 24), 8, 5.0, 17)
    b = torch.from_numpy(_rhs(cfg, cfg.sources[0])).cuda()
    z = torch.zeros_like(b)
    S = BatchSolver("cuda:0")
    for method in ("plain", "mixed"):
        xa, ia = S.solve([H, H], [b, z], method=method, tol=1e-6, inner_maxit=8)
        xb, ib = S.solve([H, H], [b, z], method=method, tol=1e-6, inner_maxit=8)
        assert ia[1].converged and ia[1].relres == 0.0 and not xa[1].any()
        assert torch.equal(xa[0], xb[0])
        assert ia[0].iters == ib[0].iters


def test_batch_autotune_keeps_results():
    import torch
    from helm3d import BatchSolver
    cfg, wn, H = _op((32, 32, 24), 8, 5.0, 18)
    b = torch.from_numpy(_rhs(cfg, cfg.sources[0])).cuda()
    S = BatchSolver("cuda:0")
    x0, i0 = S.solve([H, H], [b, 2 * b], method="mixed", tol=1e-6, inner_maxit=8)
    r = S.autotune([H, H], [b, 2 * b], inner_maxit=8, divisors=(1, 2))
    print("batch autotune", r, flush=True)
    L = H.get_launch()
    assert (L["tile_rows"], L["prefetch"], L["rows_per_thread"]) == (r["tile_rows"], r["prefetch"],
                                                                     r["rows_per_thread"])
    x1, i1 = S.solve([H, H], [b, 2 * b], method="mixed", tol=1e-6, inner_maxit=8)
    # the dot products reduced in the stencil epilogues follow the block
    # decomposition, so another shape changes rounding only
    for a, c, ia, ic in zip(x0, x1, i0, i1):
        assert ic.converged and abs(ia.iters - ic.iters) <= 2
        assert float((a - c).norm() / a.norm()) < 1e-4


def test_uneven_workload_chunked():
    """The uneven batched workload (8 subproblems with different grids and
    frequencies, 4-32 impulses each) at reduced size: every right-hand side
    reaches 1e-6 with the oracle operator, and splitting the items into
    several batches under a small workspace budget gives the same bits as one
    batch (items of a mixed batch are independent)."""
    import torch
    from helm3d import BatchSolver, Helm3d
    from helm3d.models import config_c4, extend, point_source, vec, wavenumber
    subs = config_c4(scale=0.2)
    ops, bs, meta = [], [], []
    for c in subs:
        wn = vec(wavenumber(extend(c.v, c.npml), c.freq))
        H = Helm3d(c.nt, (c.h,) * 3, c.npml6, torch.from_numpy(wn).to("cuda:0"))
        for s in c.sources:
            ops.append(H)
            bs.append(torch.from_numpy(point_source(c, s)).cuda())
            meta.append((c, wn))
    print("uneven workload:", len(subs), "grids,", len(bs), "right-hand sides", flush=True)
    S = BatchSolver("cuda:0")
    per = BatchSolver.item_bytes(ops[0], "mixed", 20)
    assert per > 41 * 16 * ops[0].N
    budget = sum(BatchSolver.item_bytes(H, "mixed", 20) for H in ops) // 3 + 1
    xs, infos, chunks = S.solve_chunked(ops, bs, method="mixed", tol=1e-6, inner_maxit=10, budget_bytes=budget)
    print("batches", [len(ch) for ch in chunks], flush=True)
    assert len(chunks) >= 3 and sorted(i for ch in chunks for i in ch) == list(range(len(bs)))
    worst = 0.0
    for (c, wn), b, x, inf in zip(meta, bs, xs, infos):
        rel = _relres(c, wn, b.cpu().numpy(), x.cpu().numpy(), 1)
        worst = max(worst, rel)
        assert inf.converged and rel < 1e-6 and abs(rel - inf.relres) < 1e-9
    print("worst relres", worst, "outer iters", min(i.iters for i in infos), "-", max(i.iters for i in infos))
    x1, i1 = S.solve(ops, bs, method="mixed", tol=1e-6, inner_maxit=10)
    for a, b_, ia, ib in zip(xs, x1, infos, i1):
        assert ia.iters == ib.iters and torch.equal(a, b_)
