"""CPU checks of the multigrid set-up pieces (no GPU): the oracle's
trilinear transfer (LinInterp1D.m / fine2coarse.m 'linear') reproduces linear
functions exactly and is a partition of unity; the product's model coarsening
(fine2coarse.m 'cubic') reproduces cubic polynomials exactly."""
import numpy as np

import oracle as O


def test_prolongation_reproduces_linear_functions():
    nf, nc = (13, 11, 9), (7, 6, 5)
    P = O.prolongation(nf, nc)
    assert P.shape == (np.prod(nf), np.prod(nc))
    assert np.allclose(P @ np.ones(P.shape[1]), 1.0, atol=1e-14)
    gc = [np.linspace(0, 1, n) for n in nc]
    gf = [np.linspace(0, 1, n) for n in nf]
    Zc, Yc, Xc = np.meshgrid(gc[2], gc[1], gc[0], indexing="ij")
    Zf, Yf, Xf = np.meshgrid(gf[2], gf[1], gf[0], indexing="ij")
    fc = (1 + 2 * Xc - 3 * Yc + 0.5 * Zc).reshape(-1)
    ff = (1 + 2 * Xf - 3 * Yf + 0.5 * Zf).reshape(-1)
    assert np.max(abs(P @ fc - ff)) < 1e-13


def test_model_coarsening_is_exact_on_cubics():
    from helm3d.mg import coarsen_model
    n, nc = (12, 10, 9), (6, 5, 5)
    g = [np.linspace(0, 1, a) for a in n]
    h = [np.linspace(0, 1, a) for a in nc]
    Z, Y, X = np.meshgrid(g[2], g[1], g[0], indexing="ij")
    Zc, Yc, Xc = np.meshgrid(h[2], h[1], h[0], indexing="ij")
    f = lambda x, y, z: 1 + x ** 3 - 2 * y ** 2 * z + x * y * z  # noqa: E731
    vc = coarsen_model(f(X, Y, Z).reshape(-1), n, nc)
    assert np.max(abs(vc - f(Xc, Yc, Zc).reshape(-1))) < 1e-12


def test_model_coarsening_needs_four_points_per_axis():
    """opLInterp1D's window starts at x1(1..4) (opLInterp1D.m:114-153), so
    the reference cannot coarsen an axis of fewer than 4 points; the
    coarsening raises there, and exactly 4 points still reproduce cubics."""
    import pytest
    from helm3d.mg import coarsen_model
    with pytest.raises(ValueError, match="at least 4"):
        coarsen_model(np.ones(3 * 5 * 5), (3, 5, 5), (2, 3, 3))
    n, nc = (4, 4, 5), (2, 2, 3)
    g = [np.linspace(0, 1, a) for a in n]
    h = [np.linspace(0, 1, a) for a in nc]
    Z, Y, X = np.meshgrid(g[2], g[1], g[0], indexing="ij")
    Zc, Yc, Xc = np.meshgrid(h[2], h[1], h[0], indexing="ij")
    f = lambda x, y, z: 2 - x ** 3 + y * z ** 2  # noqa: E731
    vc = coarsen_model(f(X, Y, Z).reshape(-1), n, nc)
    assert np.max(abs(vc - f(Xc, Yc, Zc).reshape(-1))) < 1e-12
