"""GPU Kaczmarz sweeps (27-colour order) and CGMN against the CPU oracle.

Bars: a device double sweep equals the oracle's 27-colour sweep to 1e-12
relative (forward, adjoint, complex wn; several grids incl. asymmetric PML
and a free surface); the host drop-in gateway gives the same result with the
MEX error identifiers; CGMN reaches the preconditioned tolerance, its true
residual is small, and its solution matches the oracle's CGMN and FGMRES.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    ((12, 10, 8), [[3, 2, 2], [3, 2, 2]], (10.0, 10.0, 10.0)),
    ((13, 11, 7), [[2, 3, 1], [4, 2, 3]], (10.0, 12.0, 9.0)),
    ((33, 35, 31), [[0, 4, 4], [4, 4, 4]], (10.0, 10.0, 10.0)),
]


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _inputs(n, seed, wnc):
    rng = np.random.default_rng(seed)
    N = int(np.prod(n))
    wn = rng.uniform(2e-4, 2e-3, N)
    if wnc:
        wn = wn + 1j * rng.uniform(-2e-4, 2e-4, N)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    return wn, x, b


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("variant", ["forw", "adj", "forw_wnc", "adj_wnc"])
def test_device_sweep_matches_oracle(case, variant):
    import torch
    from helm3d import Helm3d
    n, npml, h = CASES[case]
    wn, x, b = _inputs(n, 40 + case, variant.endswith("wnc"))
    print(f"sweep case={case} n={n} variant={variant}", flush=True)
    H = Helm3d(n, h, np.array(npml), torch.from_numpy(wn).cuda())
    mode = 2 if variant.startswith("adj") else 1
    for omega in (1.0, 1.5):
        y = H.kacz_sweep(torch.from_numpy(x).cuda(), torch.from_numpy(b).cuda(), mode=mode, omega=omega).cpu().numpy()
        yo = O.kacz_sweep(variant, wn, h, n, np.array(npml), omega, x, b, order="mc")
        assert relerr(y, yo) <= 1e-12
    # zero right-hand side (the preconditioned operator of CGMN), two sweeps
    y2 = H.kacz_sweep(torch.from_numpy(x).cuda(), None, mode=mode, omega=1.5, nsweeps=2).cpu().numpy()
    z = np.zeros_like(b)
    yo2 = O.kacz_sweep(variant, wn, h, n, np.array(npml), 1.5,
                       O.kacz_sweep(variant, wn, h, n, np.array(npml), 1.5, x, z), z)
    assert relerr(y2, yo2) <= 1e-12


# grids with an axis of one or two points: colours without rows on an axis
# are skipped, every row sits on the grid edge
DEGENERATE = [
    ((5, 1, 4), [[1, 1, 1], [1, 1, 1]], (10.0, 10.0, 10.0)),
    ((2, 2, 2), [[1, 1, 1], [1, 1, 1]], (10.0, 10.0, 10.0)),
    ((1, 1, 1), [[1, 1, 1], [1, 1, 1]], (10.0, 10.0, 10.0)),
    ((4, 3, 1), [[1, 1, 0], [1, 1, 1]], (10.0, 10.0, 10.0)),
]


@pytest.mark.parametrize("case", range(len(DEGENERATE)))
@pytest.mark.parametrize("variant", ["forw", "adj"])
def test_device_sweep_degenerate_grids(case, variant):
    import torch
    from helm3d import Helm3d
    n, npml, h = DEGENERATE[case]
    wn, x, b = _inputs(n, 80 + case, False)
    print(f"degenerate sweep case={case} n={n} variant={variant}", flush=True)
    H = Helm3d(n, h, np.array(npml), torch.from_numpy(wn).cuda())
    mode = 2 if variant == "adj" else 1
    y = H.kacz_sweep(torch.from_numpy(x).cuda(), torch.from_numpy(b).cuda(), mode=mode, omega=1.5).cpu().numpy()
    yo = O.kacz_sweep(variant, wn, h, n, np.array(npml), 1.5, x, b, order="mc")
    assert relerr(y, yo) <= 1e-12


@pytest.mark.parametrize("case", range(len(DEGENERATE)))
@pytest.mark.parametrize("solver", ["cgmn", "crmn"])
def test_solvers_degenerate_grids(case, solver):
    """CGMN and CRMN on grids with an axis of one or two points, forward and
    adjoint: the oracle's iteration count (1 to 10 here; on 1 x 1 x 1 the
    first step can leave an exactly zero residual) and its solution."""
    import torch
    from helm3d import Helm3d
    n, npml, h = DEGENERATE[case]
    rng = np.random.default_rng(90 + case)
    N = int(np.prod(n))
    wn = rng.uniform(2e-4, 2e-3, N)
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    H = Helm3d(n, h, np.array(npml), torch.from_numpy(wn).cuda())
    for mode, v in ((1, "forw"), (2, "adj")):
        print(f"degenerate {solver} case={case} n={n} mode={mode}", flush=True)
        fn = H.solve_cgmn if solver == "cgmn" else H.solve_crmn
        x, info = fn(torch.from_numpy(b).cuda(), mode=mode, tol=1e-8, maxit=500)
        x = x.cpu().numpy()
        sweep = lambda y, r: O.kacz_sweep(v, wn, h, n, np.array(npml), 1.5, y, r)  # noqa: E731
        xo, its, hst = (O.cgmn if solver == "cgmn" else O.crmn)(sweep, b, tol=1e-8, maxit=500)
        print(f"  gpu {info.iters} {info.quasi_relres:.1e} oracle {its} {hst[-1]:.1e}", flush=True)
        assert np.isfinite(x).all()
        assert info.converged and info.quasi_relres <= 1e-8
        assert abs(info.iters - its) <= 0.02 * its + 2
        assert relerr(x, xo) <= 1e-6
        rel = np.linalg.norm(b - O.hmvp(v, wn, h, n, np.array(npml), x)) / np.linalg.norm(b)
        assert abs(rel - info.relres) <= 1e-9 + 1e-6 * rel


def test_dropin_gateway_and_errors():
    from helm3d import Helm3dError, helm3d_kaczswp_host
    n, npml, h = CASES[1]
    wn, x, b = _inputs(n, 50, False)
    y = helm3d_kaczswp_host("forw", wn, h, n, np.array(npml), 1.5, x, b)
    assert relerr(y, O.kacz_sweep("forw", wn, h, n, np.array(npml), 1.5, x, b)) <= 1e-12
    y = helm3d_kaczswp_host("adj", wn, h, n, np.array(npml), 1.2, x.real, b)  # real x
    assert relerr(y, O.kacz_sweep("adj", wn, h, n, np.array(npml), 1.2, x.real.astype(complex), b)) <= 1e-12
    for bad, ident in (((n[0], n[1]), "Nsize"),):
        with pytest.raises(Helm3dError) as e:
            helm3d_kaczswp_host("forw", wn, h, bad, np.array(npml), 1.5, x, b)
        assert e.value.identifier.endswith("Helm3dkaczswp_mex:" + ident)
    with pytest.raises(Helm3dError) as e:
        helm3d_kaczswp_host("forw", wn, h, n, np.array(npml), 1.5, x[:-1], b)
    assert e.value.identifier.endswith("Helm3dkaczswp_mex:Xsize")
    with pytest.raises(Helm3dError) as e:
        helm3d_kaczswp_host("forw", wn, h, n, np.array(npml), 1.5, x, b[:-1])
    assert e.value.identifier.endswith("Helm3dkaczswp_mex:Bsize")


@pytest.mark.parametrize("mode", [1, 2])
def test_cgmn_matches_oracle(mode):
    import torch
    from helm3d import Helm3d
    from helm3d.models import config_small, extend, wavenumber, vec, point_source
    cfg = config_small(n=(20, 18, 16), npml=4, freq=4.0)
    wn = vec(wavenumber(extend(cfg.v, cfg.npml), cfg.freq))
    b = point_source(cfg, cfg.sources[0])
    h3, p6 = (cfg.h,) * 3, np.array(cfg.npml6)
    v = "forw" if mode == 1 else "adj"
    H = Helm3d(cfg.nt, h3, cfg.npml6, torch.from_numpy(wn).cuda())
    print("cgmn mode", mode, cfg.nt, flush=True)
    x, info = H.solve_cgmn(torch.from_numpy(b).cuda(), mode=mode, tol=1e-8, maxit=5000)
    x = x.cpu().numpy()
    sweep = lambda y, r: O.kacz_sweep(v, wn, h3, cfg.nt, p6, 1.5, y, r)  # noqa: E731
    xo, its, hst = O.cgmn(sweep, b, tol=1e-8, maxit=5000)
    print("gpu cgmn", info.iters, info.quasi_relres, info.relres, "oracle", its, hst[-1], flush=True)
    assert info.converged and info.quasi_relres <= 1e-8
    assert abs(info.iters - its) <= 0.02 * its + 2
    assert relerr(x, xo) <= 1e-6
    r = b - O.hmvp(v, wn, h3, cfg.nt, p6, x)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    assert abs(rel - info.relres) <= 1e-9 + 1e-6 * rel
    xf, _ = O.fgmres(v, wn, h3, cfg.nt, p6, b, tol=1e-10, maxinnerit=20)
    assert relerr(x, xf) <= 1e-4


@pytest.mark.parametrize("mode", [1, 2])
def test_crmn_matches_oracle(mode):
    import torch
    from helm3d import Helm3d
    from helm3d.models import config_small, extend, wavenumber, vec, point_source
    cfg = config_small(n=(20, 18, 16), npml=4, freq=4.0)
    wn = vec(wavenumber(extend(cfg.v, cfg.npml), cfg.freq))
    b = point_source(cfg, cfg.sources[0])
    h3, p6 = (cfg.h,) * 3, np.array(cfg.npml6)
    v = "forw" if mode == 1 else "adj"
    H = Helm3d(cfg.nt, h3, cfg.npml6, torch.from_numpy(wn).cuda())
    print("crmn mode", mode, cfg.nt, flush=True)
    x, info = H.solve_crmn(torch.from_numpy(b).cuda(), mode=mode, tol=1e-8, maxit=5000)
    x = x.cpu().numpy()
    sweep = lambda y, r: O.kacz_sweep(v, wn, h3, cfg.nt, p6, 1.5, y, r)  # noqa: E731
    xo, its, hst = O.crmn(sweep, b, tol=1e-8, maxit=5000)
    print("gpu crmn", info.iters, info.quasi_relres, info.relres, "oracle", its, hst[-1], flush=True)
    assert info.converged and info.quasi_relres <= 1e-8
    assert abs(info.iters - its) <= 0.02 * its + 2
    assert relerr(x, xo) <= 1e-6
    r = b - O.hmvp(v, wn, h3, cfg.nt, p6, x)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    assert abs(rel - info.relres) <= 1e-9 + 1e-6 * rel
