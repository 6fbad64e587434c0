"""CPU checks that pin the Kaczmarz sweep restatement (Helm3dkaczswp_mex.c)
in the oracle.

The reference sweep cannot be built here (its file needs MATLAB's mex.h), so
the restatement is pinned by what the mathematics fixes:
  * each row's coefficients equal that row of the matrix assembled from the
    stencil (H, or H^H for the adjoint sweep);
  * a double sweep leaves the exact solution fixed, and for 0 < omega < 2 it
    does not increase the error ||x - x*|| (each projection is a relaxed
    orthogonal projection; a wrong conjugate or normalisation breaks this);
  * the double sweep's linear part Q makes I - Q Hermitian (forward pass
    followed by its reverse), which CGMN relies on;
  * CGMN on the preconditioned system converges to the direct solution, in
    both the reference's lexicographic order and the 27-colour order used on
    the GPU (the two orders give different iterates).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
import fwi_oracle as F

N3 = (9, 8, 7)
NPML = np.array([[2, 2, 2], [2, 2, 2]])
H3 = (10.0, 10.0, 10.0)


def _wn(seed=0, wnc=False):
    rng = np.random.default_rng(seed)
    N = int(np.prod(N3))
    v = rng.uniform(1500, 3000, N)
    wn = (2 * np.pi * 8.0) ** 2 / v ** 2
    if wnc:
        wn = wn * (1 + 0.05j)
    return wn


@pytest.mark.parametrize("variant", ["forw", "adj", "forw_wnc", "adj_wnc"])
def test_rows_match_assembled_matrix(variant):
    wn = _wn(1, variant.endswith("wnc"))
    A = F.assemble(variant, wn, H3, N3, NPML).tocsr()
    for (i, j, k) in [(0, 0, 0), (4, 3, 2), (8, 7, 6), (1, 7, 3), (4, 0, 6)]:
        c, idx = O.row(variant, wn, H3, N3, NPML, i, j, k)
        p = i + N3[0] * (j + N3[1] * k)
        dense = np.zeros(A.shape[1], complex)
        dense[idx] = c
        assert np.allclose(dense, A[p].toarray().ravel(), rtol=0, atol=1e-15 * abs(c).max())


@pytest.mark.parametrize("order", ["lex", "mc"])
@pytest.mark.parametrize("variant", ["forw", "adj"])
def test_sweep_fixed_point_and_error_decrease(order, variant):
    wn = _wn(2)
    A = F.assemble(variant, wn, H3, N3, NPML).tocsc()
    rng = np.random.default_rng(3)
    b = rng.standard_normal(A.shape[0]) + 1j * rng.standard_normal(A.shape[0])
    xs = spla.splu(A).solve(b)
    y = O.kacz_sweep(variant, wn, H3, N3, NPML, 1.5, xs, b, order=order)
    assert np.linalg.norm(y - xs) <= 1e-11 * np.linalg.norm(xs)
    x = xs + 0.1 * (rng.standard_normal(xs.size) + 1j * rng.standard_normal(xs.size))
    for omega in (0.5, 1.0, 1.5, 1.9):
        e0 = np.linalg.norm(x - xs)
        x1 = O.kacz_sweep(variant, wn, H3, N3, NPML, omega, x, b, order=order)
        assert np.linalg.norm(x1 - xs) < e0


@pytest.mark.parametrize("order", ["lex", "mc"])
def test_double_sweep_gives_hermitian_system(order):
    wn = _wn(4)
    N = wn.size
    z = np.zeros(N, complex)
    rng = np.random.default_rng(5)
    u = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    v = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    Au = u - O.kacz_sweep("forw", wn, H3, N3, NPML, 1.5, u, z, order=order)
    Av = v - O.kacz_sweep("forw", wn, H3, N3, NPML, 1.5, v, z, order=order)
    lhs, rhs = np.vdot(v, Au), np.vdot(Av, u)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    assert np.vdot(u, Au).real > 0


@pytest.mark.parametrize("order", ["lex", "mc"])
def test_cgmn_converges_to_direct_solution(order):
    wn = _wn(6)
    A = F.assemble("forw", wn, H3, N3, NPML).tocsc()
    b = np.zeros(A.shape[0], complex)
    b[A.shape[0] // 2 + 5] = 1.0
    xs = spla.splu(A).solve(b)
    sweep = lambda x, r: O.kacz_sweep("forw", wn, H3, N3, NPML, 1.5, x, r, order=order)  # noqa: E731
    x, its, hst = O.cgmn(sweep, b, tol=1e-10, maxit=2000)
    assert hst[-1] <= 1e-10
    assert np.linalg.norm(x - xs) <= 1e-7 * np.linalg.norm(xs)
    print(order, "cgmn iterations", its)


@pytest.mark.parametrize("order", ["lex", "mc"])
def test_crmn_converges_to_direct_solution(order):
    """CRMN.m restatement: converges to the sparse direct solution, and its
    residual history is monotone (conjugate residuals minimise ||r'|| over
    the Krylov space of the Hermitian A')."""
    wn = _wn(6)
    A = F.assemble("forw", wn, H3, N3, NPML).tocsc()
    b = np.zeros(A.shape[0], complex)
    b[A.shape[0] // 2 + 5] = 1.0
    xs = spla.splu(A).solve(b)
    sweep = lambda x, r: O.kacz_sweep("forw", wn, H3, N3, NPML, 1.5, x, r, order=order)  # noqa: E731
    x, its, hst = O.crmn(sweep, b, tol=1e-10, maxit=2000)
    assert hst[-1] <= 1e-10
    assert np.linalg.norm(x - xs) <= 1e-7 * np.linalg.norm(xs)
    assert np.all(np.diff(hst) <= 1e-12 * hst[0])
    print(order, "crmn iterations", its)
