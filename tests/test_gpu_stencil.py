"""GPU parity of the 27-point stencil against the CPU oracle (all six MEX variants).

Bar: <= 1e-12 relative (||y_gpu - y_oracle|| / ||y_oracle||) for complex128,
<= 2e-6 for the complex64 kernel; adjoint identities to ~1e-14.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

# (n incl. PML, npml 2x3 [lo;hi], h): shapes cover several x tiles (64 wide),
# partial y tiles (4 rows), several z chunks, asymmetric PML and a free surface.
CASES = [
    ((12, 10, 8), [[3, 2, 2], [3, 2, 2]], (10.0, 10.0, 10.0)),
    ((9, 11, 7), [[2, 3, 1], [4, 2, 3]], (10.0, 12.0, 9.0)),
    ((70, 9, 13), [[5, 2, 3], [5, 2, 3]], (12.5, 12.5, 12.5)),
    ((130, 7, 40), [[10, 2, 6], [10, 2, 6]], (10.0, 10.0, 10.0)),
    ((33, 35, 31), [[0, 4, 4], [4, 4, 4]], (10.0, 10.0, 10.0)),
    # whole 64-wide waves outside the x/y PML (interior path) next to PML waves
    ((200, 14, 10), [[4, 2, 2], [5, 3, 2]], (10.0, 11.0, 12.0)),
]
VARIANTS = ("forw", "adj", "forw_deriv", "adj_deriv", "forw_wnc", "adj_wnc")


def _inputs(n, seed, wnc):
    rng = np.random.default_rng(seed)
    N = int(np.prod(n))
    wn = rng.uniform(2e-4, 2e-3, N)
    if wnc:
        wn = wn + 1j * rng.uniform(-2e-4, 2e-4, N)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    return wn, x


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("variant", VARIANTS)
def test_dropin_matches_oracle(case, variant):
    from helm3d import helm3d_mvp_host
    n, npml, h = CASES[case]
    wn, x = _inputs(n, 100 + case, variant.endswith("wnc"))
    print(f"dropin case={case} n={n} variant={variant}", flush=True)
    y = helm3d_mvp_host(variant, wn, h, n, np.array(npml), x)
    yo = O.hmvp(variant, wn, h, n, np.array(npml), x)
    assert relerr(y, yo) <= 1e-12
    if O.ref_available():
        yr = O.hmvp(variant, wn, h, n, np.array(npml), x, use_ref=True)
        assert relerr(y, yr) <= 1e-12


def test_dropin_real_x_and_errors():
    from helm3d import helm3d_mvp_host, Helm3dError
    n, npml, h = CASES[0]
    wn, x = _inputs(n, 7, False)
    y = helm3d_mvp_host("forw", wn, h, n, np.array(npml), x.real)
    yo = O.hmvp("forw", wn, h, n, np.array(npml), x.real.astype(complex))
    assert relerr(y, yo) <= 1e-12
    with pytest.raises(Helm3dError) as e:
        helm3d_mvp_host("forw", wn, h, (12, 10), np.array(npml), x)
    assert e.value.identifier.endswith("Helm3dmvp_mex:Nsize")
    with pytest.raises(Helm3dError) as e:
        helm3d_mvp_host("forw", wn, h, n, np.array(npml), x[:-1])
    assert e.value.identifier.endswith("Helm3dkaczswp_mex:Xsize")


@pytest.mark.parametrize("case", [1, 3, 4, 5])
@pytest.mark.parametrize("wnc", [False, True])
def test_device_apply_and_adjoint(case, wnc):
    import torch
    from helm3d import Helm3d
    n, npml, h = CASES[case]
    wn, x = _inputs(n, 300 + case, wnc)
    dev = torch.device("cuda:0")
    print(f"device case={case} n={n} wnc={wnc}", flush=True)
    H = Helm3d(n, h, np.array(npml), torch.from_numpy(wn).to(dev))
    xt = torch.from_numpy(x).to(dev)
    yf = H.apply(xt).cpu().numpy()
    ya = H.apply(xt, adjoint=True).cpu().numpy()
    vf = "forw_wnc" if wnc else "forw"
    va = "adj_wnc" if wnc else "adj"
    assert relerr(yf, O.hmvp(vf, wn, h, n, np.array(npml), x)) <= 1e-12
    assert relerr(ya, O.hmvp(va, wn, h, n, np.array(npml), x)) <= 1e-12
    # <H x, z> = <x, H^H z>
    rng = np.random.default_rng(9)
    z = rng.standard_normal(x.size) + 1j * rng.standard_normal(x.size)
    Haz = H.apply(torch.from_numpy(z).to(dev), adjoint=True).cpu().numpy()
    lhs = np.vdot(z, yf)
    rhs = np.vdot(Haz, x)
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    # complex64 kernel
    y32 = H.apply(xt.to(torch.complex64)).cpu().numpy()
    assert relerr(y32, yf) <= 2e-6


def test_device_deriv():
    import torch
    from helm3d import Helm3d
    n, npml, h = CASES[2]
    wn, x = _inputs(n, 11, False)
    rng = np.random.default_rng(12)
    dwn = rng.uniform(-1e-3, 1e-3, x.size)
    dev = torch.device("cuda:0")
    H = Helm3d(n, h, np.array(npml), torch.from_numpy(wn).to(dev))
    xt = torch.from_numpy(x).to(dev)
    dt = torch.from_numpy(dwn).to(dev)
    yf = H.apply_deriv(dt, xt).cpu().numpy()
    ya = H.apply_deriv(dt, xt, adjoint=True).cpu().numpy()
    assert relerr(yf, O.hmvp("forw_deriv", dwn, h, n, np.array(npml), x)) <= 1e-12
    assert relerr(ya, O.hmvp("adj_deriv", dwn, h, n, np.array(npml), x)) <= 1e-12


# Degenerate grids: an axis of one or two points (every column or row on the
# grid edge, one-plane z chunks); the oracle agrees with Helm3d_27pt.h compiled
# in place on all of them.
DEGENERATE = [
    ((5, 1, 4), [[1, 1, 1], [1, 1, 1]], (10.0, 10.0, 10.0)),
    ((1, 6, 5), [[1, 2, 1], [1, 2, 1]], (10.0, 12.0, 9.0)),
    ((4, 3, 1), [[1, 1, 0], [1, 1, 1]], (10.0, 10.0, 10.0)),
    ((2, 2, 2), [[1, 1, 1], [1, 1, 1]], (10.0, 10.0, 10.0)),
    ((1, 1, 1), [[1, 1, 1], [1, 1, 1]], (10.0, 10.0, 10.0)),
    ((70, 2, 3), [[5, 1, 1], [5, 1, 1]], (12.5, 12.5, 12.5)),
]


@pytest.mark.parametrize("case", range(len(DEGENERATE)))
def test_degenerate_grids(case):
    """Drop-in (all six variants) and device products on grids with an axis
    of one or two points, against the oracle and the compiled reference, and
    bit-identical across launch shapes including one-plane z chunks."""
    import torch
    from helm3d import Helm3d, helm3d_mvp_host
    n, npml, h = DEGENERATE[case]
    print(f"degenerate case={case} n={n}", flush=True)
    for variant in VARIANTS:
        wn, x = _inputs(n, 700 + case, variant.endswith("wnc"))
        y = helm3d_mvp_host(variant, wn, h, n, np.array(npml), x)
        assert relerr(y, O.hmvp(variant, wn, h, n, np.array(npml), x)) <= 1e-12, variant
        if O.ref_available():
            assert relerr(y, O.hmvp(variant, wn, h, n, np.array(npml), x, use_ref=True)) <= 1e-12, variant
    wn, x = _inputs(n, 750 + case, False)
    dev = torch.device("cuda:0")
    H = Helm3d(n, h, np.array(npml), torch.from_numpy(wn).to(dev))
    xt = torch.from_numpy(x).to(dev)
    ref = [H.apply(xt), H.apply(xt, adjoint=True), H.apply(xt.to(torch.complex64))]
    assert relerr(ref[0].cpu().numpy(), O.hmvp("forw", wn, h, n, np.array(npml), x)) <= 1e-12
    assert relerr(ref[1].cpu().numpy(), O.hmvp("adj", wn, h, n, np.array(npml), x)) <= 1e-12
    for shape in [(1, 4, 1, 1), (1, 8, 2, 2), (2, 8, 4, 1), (0, 8, 1, 2)]:
        H.set_launch(*shape)
        got = [H.apply(xt), H.apply(xt, adjoint=True), H.apply(xt.to(torch.complex64))]
        for k, (a, b) in enumerate(zip(got, ref)):
            assert torch.equal(a, b), (shape, k)


SHAPES = [(0, 4, 1, 1), (35, 4, 2, 1), (70, 8, 1, 1), (17, 8, 2, 1), (23, 4, 3, 1), (9, 8, 4, 1), (70, 4, 4, 1),
          (0, 8, 1, 2), (70, 8, 2, 2), (17, 8, 3, 2), (9, 8, 4, 2),
          (1, 4, 2, 1), (1, 8, 4, 1), (2, 4, 3, 1), (2, 8, 2, 2), (4, 4, 1, 1), (1, 8, 1, 2)]


@pytest.mark.parametrize("shape", SHAPES)
def test_launch_shapes_identical(shape):
    """Every launch shape (z chunk, tile rows, prefetch, rows per thread) gives
    bit-identical results, interior and PML waves: forward and adjoint with
    real and complex wn, the derivative operator, complex64, and the gradient
    epilogue."""
    import torch
    from helm3d import Helm3d
    n, npml, h = CASES[5]
    wn, x = _inputs(n, 55, False)
    wnc, _ = _inputs(n, 56, True)
    dev = torch.device("cuda:0")
    H = Helm3d(n, h, np.array(npml), torch.from_numpy(wn).to(dev))
    Hc = Helm3d(n, h, np.array(npml), torch.from_numpy(wnc).to(dev))
    xt = torch.from_numpy(x).to(dev)
    dwn = torch.from_numpy(np.random.default_rng(57).uniform(0.5, 2.0, wn.size)).to(dev)
    U = torch.flip(xt, [0]).contiguous()

    def products():
        g = torch.zeros(wn.size, dtype=torch.float64, device=dev)
        H.grad_accum(U, xt, g, scale=0.7, grid="pml", dwn=dwn)
        gp = torch.ones(int(np.prod(H.n_phys)), dtype=torch.float64, device=dev)
        H.grad_accum(U, xt, gp, scale=1.3, grid="phys")
        return [H.apply(xt), H.apply(xt, adjoint=True), Hc.apply(xt), Hc.apply(xt, adjoint=True),
                H.apply_deriv(dwn, xt), H.apply_deriv(dwn, xt, adjoint=True), H.apply(xt.to(torch.complex64)),
                H.apply(xt.to(torch.complex64), adjoint=True), g, gp]

    ref = [t.clone() for t in products()]
    H.set_launch(*shape)
    Hc.set_launch(*shape)
    for k, (a, b) in enumerate(zip(products(), ref)):
        assert torch.equal(a, b), (k, float((a - b).abs().max() / b.abs().max()))
    assert relerr(ref[0].cpu().numpy(), O.hmvp("forw", wn, h, n, np.array(npml), x)) <= 1e-12


def test_solve_with_two_rows_per_thread():
    """A solve on the 64 x 8, two-rows-per-thread shape: the products are
    bit-identical and only the in-block order of the norm partials changes,
    so the iteration count is the same and x agrees to rounding."""
    import torch
    from helm3d import Helm3d
    from helm3d.models import config_small, extend, point_source, vec, wavenumber
    cfg = config_small()
    wn = vec(wavenumber(extend(cfg.v, cfg.npml), cfg.freq))
    b = torch.from_numpy(point_source(cfg, cfg.sources[0])).cuda()
    H = Helm3d(cfg.nt, (cfg.h,) * 3, cfg.npml6, torch.from_numpy(wn).cuda())
    x1, i1 = H.solve(b, tol=1e-8)
    H.set_launch(0, 8, 2, 2)
    x2, i2 = H.solve(b, tol=1e-8)
    assert i1.converged and i2.converged and i1.iters == i2.iters
    assert float((x1 - x2).norm() / x1.norm()) < 1e-9


def test_c2_full_size_matvec_parity():
    """The benchmark operator (240x240x140, PML 20) at full size, fwd and adj."""
    import torch
    from helm3d import Helm3d
    from helm3d.models import config_c2, extend, wavenumber, vec
    cfg = config_c2()
    wn = vec(wavenumber(extend(cfg.v, cfg.npml), cfg.freq))
    rng = np.random.default_rng(2026)
    x = rng.standard_normal(wn.size) + 1j * rng.standard_normal(wn.size)
    dev = torch.device("cuda:0")
    H = Helm3d(cfg.nt, (cfg.h,) * 3, cfg.npml6, torch.from_numpy(wn).to(dev))
    xt = torch.from_numpy(x).to(dev)
    for adj, v in ((False, "forw"), (True, "adj")):
        y = H.apply(xt, adjoint=adj).cpu().numpy()
        yo = O.hmvp(v, wn, (cfg.h,) * 3, cfg.nt, np.array(cfg.npml6), x)
        assert relerr(y, yo) <= 1e-12, v


def test_autotune_keeps_results():
    """The autotuner picks a valid shape; products are unchanged by it."""
    import torch
    from helm3d import Helm3d
    n, npml, h = CASES[5]
    wn, x = _inputs(n, 77, False)
    dev = torch.device("cuda:0")
    H = Helm3d(n, h, np.array(npml), torch.from_numpy(wn).to(dev))
    xt = torch.from_numpy(x).to(dev)
    y0 = H.apply(xt).clone()
    a0 = H.apply(xt, adjoint=True).clone()
    s0 = H.apply(xt.to(torch.complex64)).clone()
    for mode, prec in ((1, "c128"), (2, "c128"), (1, "c64")):
        r = H.autotune(mode=mode, precision=prec, reps=3)
        print("autotune", mode, prec, r, flush=True)
        assert r["tile_rows"] in (4, 8) and r["prefetch"] in (1, 2, 3, 4) and 1 <= r["kchunk"] <= n[2] and r["ms"] > 0
        assert r["rows_per_thread"] == 1 or (r["rows_per_thread"] == 2 and r["tile_rows"] == 8)
        L = H.get_launch()
        keys = ("kchunk", "tile_rows", "prefetch", "rows_per_thread")
        assert tuple(L[q] for q in keys) == tuple(r[q] for q in keys)
        assert torch.equal(H.apply(xt), y0)
        assert torch.equal(H.apply(xt, adjoint=True), a0)
        assert torch.equal(H.apply(xt.to(torch.complex64)), s0)
