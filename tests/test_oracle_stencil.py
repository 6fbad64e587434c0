"""CPU checks that pin the stencil oracle (no GPU).

The reference ships no golden vectors or fixtures for this path (its 3D tests
are invariant-based), so the oracle is pinned by
  * the reference stencil header compiled in place (oracle/_ref), all six
    variants, symmetric / asymmetric / free-surface PML;
  * the committed golden fixtures written from oracle/_ref
    (tests/golden/make_golden.py), which also travel to machines without the
    reference tree;
  * identities that a plausible mistake breaks: <Hx,z> = <x,H^H z> for
    (forw, adj), (forw_wnc, adj_wnc), (forw_deriv, adj_deriv); linearity of
    the DERIV stencil in dwn; H(wn) = H(0) + dH(wn) for real wn (the split of
    stiffness and mass); the interior closed form of a plane wave.
"""
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "stencil_golden.npz")
VARIANTS = ("forw", "adj", "forw_deriv", "adj_deriv", "forw_wnc", "adj_wnc")
CASES = [
    ((12, 10, 8), np.array([[3, 2, 2], [3, 2, 2]]), (10.0, 10.0, 10.0)),
    ((9, 11, 7), np.array([[2, 3, 1], [4, 2, 3]]), (10.0, 12.0, 9.0)),
    ((10, 10, 10), np.array([[0, 2, 2], [3, 2, 2]]), (10.0, 10.0, 10.0)),
]


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    N = int(np.prod(n))
    wn = rng.uniform(2e-4, 2e-3, N)
    wnc = wn + 1j * rng.uniform(-2e-4, 2e-4, N)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    z = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    return wn, wnc, x, z


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("variant", VARIANTS)
def test_oracle_matches_reference_build(case, variant):
    n, npml, h = CASES[case]
    wn, wnc, x, _ = _inputs(n, case)
    w = wnc if variant.endswith("wnc") else wn
    a = O.hmvp(variant, w, h, n, npml, x)
    b = O.hmvp(variant, w, h, n, npml, x, use_ref=True)
    assert relerr(a, b) <= 1e-14


def test_oracle_matches_golden_fixtures():
    g = np.load(GOLDEN, allow_pickle=False)
    ncase = int(g["ncase"])
    assert ncase >= 3
    for c in range(ncase):
        n = tuple(int(v) for v in g[f"n{c}"])
        npml = g[f"npml{c}"]
        h = tuple(float(v) for v in g[f"h{c}"])
        x = g[f"x{c}"]
        for v in VARIANTS:
            w = g[f"wnc{c}"] if v.endswith("wnc") else g[f"wn{c}"]
            y = O.hmvp(v, w, h, n, npml, x)
            assert relerr(y, g[f"y_{v}{c}"]) <= 1e-14, (c, v)


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("pair", [("forw", "adj"), ("forw_wnc", "adj_wnc"), ("forw_deriv", "adj_deriv")])
def test_dot_test(case, pair):
    n, npml, h = CASES[case]
    wn, wnc, x, z = _inputs(n, 10 + case)
    w = wnc if pair[0].endswith("wnc") else wn
    Hx = O.hmvp(pair[0], w, h, n, npml, x)
    Haz = O.hmvp(pair[1], w, h, n, npml, z)
    lhs, rhs = np.vdot(z, Hx), np.vdot(Haz, x)
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def test_deriv_linear_and_mass_split():
    n, npml, h = CASES[1]
    wn, _, x, _ = _inputs(n, 3)
    rng = np.random.default_rng(4)
    d1 = rng.standard_normal(wn.size)
    d2 = rng.standard_normal(wn.size)
    a = O.hmvp("forw_deriv", 2.0 * d1 - 3.0 * d2, h, n, npml, x)
    b = 2.0 * O.hmvp("forw_deriv", d1, h, n, npml, x) - 3.0 * O.hmvp("forw_deriv", d2, h, n, npml, x)
    assert relerr(a, b) <= 1e-14
    # H(wn) x = H(0) x + dH(wn) x : the stiffness part does not depend on wn
    full = O.hmvp("forw", wn, h, n, npml, x)
    stiff = O.hmvp("forw", np.zeros_like(wn), h, n, npml, x)
    mass = O.hmvp("forw_deriv", wn, h, n, npml, x)
    assert relerr(stiff + mass, full) <= 1e-14


def test_interior_plane_wave_closed_form():
    """Away from the PML and boundary, xi_lo = xi_hi = 1 and a plane wave
    exp(i k.x) is an eigenvector of the constant-coefficient 27-point row:
    (Hx)[p] = lambda x[p] with lambda computed independently from the
    constants (mixed-grid symbol of Operto et al. 2007)."""
    n = (20, 20, 20)
    npml = np.array([[2, 2, 2], [2, 2, 2]])
    h = (10.0, 11.0, 9.0)
    kv = np.array([0.21, -0.13, 0.17])
    I, J, K = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    ph = np.exp(1j * (kv[0] * I + kv[1] * J + kv[2] * K))
    x = ph.reshape(-1, order="F")
    wn0 = 1.3e-3
    wn = np.full(x.size, wn0)
    y = O.hmvp("forw", wn, h, n, npml, x).reshape(n, order="F")
    c = O.hconst(h)
    cx, cy, cz = np.cos(kv)
    # stiffness symbol with xi = 2, xi_lo = xi_hi = 1
    lam = -(c["cx"] * 2 + c["cy"] * 2 + c["cz"] * 2)
    lam += 2 * cx * (c["cx"] + 2 * c["xz"] + 2 * c["xy"])
    lam += 2 * cy * (c["cy"] + 2 * c["yz"] + 2 * c["xy"])
    lam += 2 * cz * (c["cz"] + 2 * c["yz"] + 2 * c["xz"])
    lam += 4 * cx * cy * (-2 * c["xy"] + 2 * c["w3a2"])
    lam += 4 * cx * cz * (-2 * c["xz"] + 2 * c["w3a2"])
    lam += 4 * cy * cz * (-2 * c["yz"] + 2 * c["w3a2"])
    lam += 8 * cx * cy * cz * (-3 * c["w3a2"])
    # mass symbol
    WM2, WM3, WM4 = 0.075168750000000, 0.004373916666667, 5.690375e-07
    lam += wn0 * (c["wn_coef"] - WM2 * 2 * (cx + cy + cz) - WM3 * 4 * (cx * cy + cx * cz + cy * cz)
                  - WM4 * 8 * cx * cy * cz)
    core = (slice(4, 16),) * 3
    assert np.allclose(y[core], lam * ph[core], rtol=1e-12, atol=1e-12 * abs(lam))


def test_pml_tables_profile():
    """xi = 2 in the interior; the lo/hi profiles are mirror images for a
    symmetric PML (x -> n+1-x), and xi_lo(x) = 2/(...)(closed form at x=1)."""
    n, p = 30, 6
    lo, hi = O.pml_axis(n, p, p)
    interior = slice(p + 2, n - p)
    assert np.allclose(lo[interior], 1.0) and np.allclose(hi[interior], 1.0)
    # symmetric PML: xi_lo at index i mirrors xi_hi at index n-1-i
    for i in range(n):
        assert abs(lo[i + 1] - hi[n - 1 - i + 1]) < 1e-12
    g = lambda x: np.cos(np.pi * (x - 1) * n / (2 * (n + 1) * p))  # noqa: E731
    expect = 2 / ((2 - g(2) * (g(2) + g(1))) + 1j * (3 * g(2) + g(1)))
    assert abs(lo[1] - expect) < 1e-15
