"""CPU checks that pin the FWI restatement in oracle/fwi_oracle.py (no GPU).

MATLAB cannot run PDEfunc.m here, so the oracle is pinned by what the
mathematics fixes:
  * the gradient of OBJ equals central finite differences of the objective
    (exact sparse solves) on seeded directions, for s2/m2 and m/s units --
    a dropped conj, a wrong sign or a forward-instead-of-adjoint stencil fails;
  * with comp2phys = Ppad' (cut_pml) the gradient is the Pext' gradient
    restricted to points whose replicas lie inside the physical box;
  * JACOB_FORW / JACOB_ADJ satisfy Re<J dm, dd> = <dm, J' dd>, and HESS_GN
    equals J'(J dm);
  * the sinc/Kaiser interpolation is a unit impulse for points on the grid and
    reconstructs a band-limited signal between grid points;
  * Pext' is the transpose of Pext (dot identity) and Pext is edge padding.
Product host code (helm3d.fwi set-up helpers) is compared with the oracle
where it has no GPU part.
"""
import numpy as np
import pytest

import oracle as O
import fwi_oracle as F

N_PHYS = (9, 8, 7)
D = (50.0, 50.0, 50.0)
NPML = np.array([[3, 3, 3], [3, 3, 3]])


def _problem(unit="s2/m2", cut_pml=False, freq=(5.0, 6.0)):
    o = (0.0, 0.0, 0.0)
    src = (np.array([100.0, 300.0]), np.array([150.0]), np.array([50.0]))
    rec = (np.array([50.0, 200.0, 350.0]), np.array([100.0, 250.0]), np.array([100.0]))
    return F.Problem(o, D, N_PHYS, freq, src, rec, NPML, unit=unit, cut_pml=cut_pml)


def _models(unit, seed=3):
    rng = np.random.default_rng(seed)
    n = int(np.prod(N_PHYS))
    v_true = 2000.0 + 300.0 * rng.standard_normal(n).clip(-2, 2)
    v0 = np.full(n, 2000.0)
    if unit == "m/s":
        return v_true, v0
    return 1.0 / v_true ** 2, 1.0 / v0 ** 2


def test_assemble_matches_stencil():
    rng = np.random.default_rng(1)
    n = (7, 6, 5)
    N = int(np.prod(n))
    wn = rng.uniform(1e-4, 1e-3, N)
    for var in ("forw", "adj", "forw_deriv"):
        A = F.assemble(var, wn, (10.0, 10.0, 10.0), n, NPML[:, :] * 0 + 2)
        x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        y = O.hmvp(var, wn, (10.0, 10.0, 10.0), n, NPML * 0 + 2, x)
        assert np.linalg.norm(A @ x - y) <= 1e-14 * np.linalg.norm(y)


def test_sinc_on_grid_is_impulse():
    xout = 25.0 * np.arange(-3, 12)
    W = F.sinc_weights(np.array([0.0, 100.0, 175.0]), xout)
    for c, x in enumerate((0.0, 100.0, 175.0)):
        e = np.zeros(xout.size)
        e[np.argmin(abs(xout - x))] = 1.0
        assert np.allclose(W[:, c], e, atol=1e-15)


def test_sinc_reconstructs_bandlimited_signal():
    dx = 10.0
    xin = dx * np.arange(0, 60)
    pts = np.array([203.7, 288.1, 311.9])
    W = F.sinc_weights(pts, xin)           # grid -> points is W'
    for kk in (0.05, 0.12):
        s = np.cos(2 * np.pi * kk * xin / dx)
        rec = W.T @ s
        assert np.max(abs(rec - np.cos(2 * np.pi * kk * pts / dx))) < 2e-2
    # the windowed kernel at half-integer offsets (closed form)
    d = 0.5
    val = np.i0(6.31 * np.sqrt(1 - (d / 4) ** 2)) / np.i0(6.31) * np.sin(np.pi * d) / (np.pi * d)
    W1 = F.sinc_weights(np.array([5.0]), dx * np.arange(3))
    assert abs(W1[0, 0] - val) < 1e-15 and abs(W1[1, 0] - val) < 1e-15


def test_product_interp_matches_oracle():
    from helm3d.fwi import interp_1d
    xin = np.array([12.0, 37.5, 100.0, 181.25])
    xout = -75.0 + 25.0 * np.arange(20)
    W = F.sinc_weights(xin, xout)
    Wp = interp_1d(xin, xout)
    assert np.max(abs(W - Wp)) <= 1e-13


def test_extension_pair():
    rng = np.random.default_rng(4)
    P = F.extension(N_PHYS, [[2, 1, 3], [1, 2, 0]])
    x = rng.standard_normal(P.shape[1])
    y = rng.standard_normal(P.shape[0])
    assert abs(np.dot(P @ x, y) - np.dot(x, P.T @ y)) <= 1e-12 * abs(np.dot(x, P.T @ y))
    X = x.reshape(N_PHYS[2], N_PHYS[1], N_PHYS[0])
    ref = np.pad(X, ((3, 0), (1, 2), (2, 1)), mode="edge").reshape(-1)
    assert np.array_equal(P @ x, ref)


@pytest.mark.parametrize("unit", ["s2/m2", "m/s"])
def test_gradient_matches_finite_differences(unit):
    P = _problem(unit=unit, cut_pml=False)
    m_true, m0 = _models(unit)
    Dobs = P.forward(m_true)
    f0, g = P.objective(m0, Dobs)
    rng = np.random.default_rng(7)
    scale = np.abs(m0).max()
    for _ in range(3):
        dm = rng.standard_normal(m0.size) * scale
        eps = 1e-5      # central differences: error O(eps^2), about 3e-8 relative here
        fp, _ = P.objective(m0 + eps * dm, Dobs)
        fm, _ = P.objective(m0 - eps * dm, Dobs)
        fd = (fp - fm) / (2 * eps)
        assert abs(fd - np.dot(g, dm)) <= 1e-6 * abs(fd), (fd, np.dot(g, dm))
    # second-order Taylor remainder: f(m + t dm) - f(m) - t g.dm = O(t^2)
    dm = rng.standard_normal(m0.size) * scale
    r = []
    for t in (1e-3, 5e-4):
        ft, _ = P.objective(m0 + t * dm, Dobs)
        r.append(abs(ft - f0 - t * np.dot(g, dm)))
    assert 3.0 < r[0] / r[1] < 5.0


def test_gradient_cut_pml_is_interior_restriction():
    unit = "s2/m2"
    m_true, m0 = _models(unit)
    Pe = _problem(unit=unit, cut_pml=False)
    Pc = _problem(unit=unit, cut_pml=True)
    Dobs = Pe.forward(m_true)
    _, ge = Pe.objective(m0, Dobs)
    _, gc = Pc.objective(m0, Dobs)
    G_e = ge.reshape(N_PHYS[::-1])
    G_c = gc.reshape(N_PHYS[::-1])
    inner = (slice(1, -1),) * 3
    assert np.allclose(G_c[inner], G_e[inner], rtol=1e-12, atol=1e-12 * abs(ge).max())
    # on the faces the truncated gradient misses the PML replicas
    assert not np.allclose(G_c, G_e, rtol=1e-3)


def test_jacobian_adjoint_pair_and_gauss_newton():
    unit = "s2/m2"
    P = _problem(unit=unit, cut_pml=False, freq=(5.0,))
    _, m0 = _models(unit)
    rng = np.random.default_rng(11)
    dm = rng.standard_normal(m0.size) * m0.max()
    Jdm = P.jacob_forw(m0, dm)
    dd = rng.standard_normal(Jdm.shape) + 1j * rng.standard_normal(Jdm.shape)
    lhs = np.real(np.vdot(dd, Jdm))
    rhs = np.dot(dm, P.jacob_adj(m0, dd))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    hgn = P.hess_gn(m0, dm)
    assert np.linalg.norm(hgn - P.jacob_adj(m0, Jdm)) <= 1e-10 * np.linalg.norm(hgn)


def test_cubic_interpolation_matches_reference_self_test():
    """opLInterp1D.m:58-94 checks its own operator: exact on cubics (zero
    outside the grid) and a consistent adjoint; on-grid points select."""
    from helm3d.fwi import lagrange_cubic, interp_1d
    rng = np.random.default_rng(2)
    x1 = np.sort(rng.uniform(0, 100, 12))
    x2 = np.sort(rng.uniform(-10, 110, 40))
    A = lagrange_cubic(x1, x2)
    f1 = x1 ** 3
    f2 = x2 ** 3
    f2[(x2 < x1[0]) | (x2 > x1[-1])] = 0
    assert np.linalg.norm(f2 - A @ f1) / np.linalg.norm(f2) < 1e-10
    g1 = rng.standard_normal(x1.size)
    g2 = rng.standard_normal(x2.size)
    assert abs(1 - ((A @ g1) @ g2) / (g1 @ (A.T @ g2))) < 1e-10
    # on-grid points reduce to a selection
    S = interp_1d(x1, x1[[0, 3, 7, 11]], "cubic")
    assert np.allclose(S, np.eye(12)[[0, 3, 7, 11]], atol=1e-12)
    with pytest.raises(ValueError):
        interp_1d(x1, x2, "linear-ish")


def test_freq_data_container():
    import torch
    from helm3d.fwi import FWIFreqData, FWIModel
    model = FWIModel(o=(0, 0, 0), d=(10, 10, 10), n=(8, 8, 8), freq=[3.0, 5.0], xsrc=[10.0, 20.0], ysrc=[10.0],
                     zsrc=[10.0], xrec=[0.0, 10.0, 20.0], yrec=[0.0], zrec=[0.0])
    D = np.arange(3 * 2 * 2) + 1j * np.arange(12)
    F = FWIFreqData(D, model)
    assert F.localData.shape == (3, 4)
    assert np.array_equal(F.getData(2).numpy(), D[6:9])  # column-major reshape(D, nrec, [])
    assert np.array_equal(FWIFreqData(D.reshape(3, 4, order="F"), model).getData(2).numpy(), D[6:9])
    F.setData([1], np.zeros(3))
    assert not F.getData(1).any()
    assert F.freqgrid == [3.0, 5.0]


@pytest.mark.parametrize("unit", ["s2/m2", "m/s"])
def test_hess_equals_gauss_newton_at_zero_residual(unit):
    """PDEfunc(HESS) at data the model fits (dphi = 0, hence V = 0): only the
    T(U)' dV term survives, which is PDEfunc(HESS_GN)."""
    P = _problem(unit=unit, cut_pml=False, freq=(5.0,))
    m_true, m0 = _models(unit)
    D0 = P.forward(m0)
    rng = np.random.default_rng(31)
    dm = rng.standard_normal(m0.size) * abs(m0).max() * 1e-2
    h = P.hess(m0, dm, D0)
    hg = P.hess_gn(m0, dm)
    assert np.linalg.norm(h - hg) <= 1e-10 * np.linalg.norm(hg)


@pytest.mark.parametrize("cut_pml", [False, True])
def test_hess_matches_gradient_differences_for_constant_dm(cut_pml):
    """For squared slowness (dwn constant, ddwn = 0) and a constant dm the
    reference's HESS is the exact derivative of the gradient (T(V) dm =
    dH (dm .* V) commutes with dm): central differences of the gradient with
    nonzero residuals agree with it."""
    P = _problem(unit="s2/m2", cut_pml=cut_pml, freq=(5.0, 6.0))
    m_true, m0 = _models("s2/m2")
    D = P.forward(m_true)
    dm = np.full(m0.size, 0.05 * m0.mean())
    h = P.hess(m0, dm, D)
    eps = 1e-3
    _, gp = P.objective(m0 + eps * dm, D)
    _, gm = P.objective(m0 - eps * dm, D)
    fd = (gp - gm) / (2 * eps)
    print("hess vs gradient differences", np.linalg.norm(h - fd) / np.linalg.norm(fd))
    assert np.linalg.norm(h - fd) <= 1e-5 * np.linalg.norm(fd)
    # and it differs from Gauss-Newton by the residual terms
    hg = P.hess_gn(m0, dm)
    assert np.linalg.norm(h - hg) > 1e-3 * np.linalg.norm(hg)
