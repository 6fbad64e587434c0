"""GPU parity of the FWI path (helm3d.fwi + the HIP set-up, sampling and
gradient kernels) against the CPU oracle (oracle/fwi_oracle.py, exact solves).

Bars: set-up kernels (param2wavenum + extension, its adjoint, the sparse
operators, the gradient stencil) <= 1e-13 relative; objective, gradient and
the PDEfunc modes with solves at tol 1e-12 <= 1e-8 relative; central finite
differences on 20 seeded directions (solves at tol 1e-10) <= 1e-4 relative.
"""
import numpy as np
import pytest

import oracle as O
import fwi_oracle as F

pytestmark = pytest.mark.gpu

N_PHYS = (9, 8, 7)
D = (50.0, 50.0, 50.0)
NPML = np.array([[3, 3, 3], [3, 3, 3]])
SRC = (np.array([100.0, 300.0]), np.array([150.0]), np.array([50.0]))
REC = (np.array([50.0, 200.0, 350.0]), np.array([100.0, 250.0]), np.array([100.0]))


def relerr(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _models(unit, seed=3, n_phys=N_PHYS):
    rng = np.random.default_rng(seed)
    n = int(np.prod(n_phys))
    v_true = 2000.0 + 300.0 * rng.standard_normal(n).clip(-2, 2)
    v0 = np.full(n, 2000.0) + 50.0 * rng.standard_normal(n).clip(-2, 2)
    if unit == "m/s":
        return v_true, v0
    return 1.0 / v_true ** 2, 1.0 / v0 ** 2


def _pair(unit="s2/m2", cut_pml=True, freq=(5.0, 6.0), tol=1e-12, solver="fgmres"):
    import torch
    from helm3d.fwi import FWIModel, FWIOpts, FWIProblem
    model = FWIModel(o=(0.0, 0.0, 0.0), d=D, n=N_PHYS, freq=list(freq), xsrc=SRC[0], ysrc=SRC[1], zsrc=SRC[2],
                     xrec=REC[0], yrec=REC[1], zrec=REC[2], f0=10.0, t0=0.0, unit=unit)
    opts = FWIOpts(pml=3, pml_max=10, cut_pml=cut_pml, tol=tol, maxit=100000, solver=solver)
    gpu = FWIProblem(model, opts, device=torch.device("cuda:0"))
    ora = F.Problem((0.0, 0.0, 0.0), D, N_PHYS, freq, SRC, REC, NPML, unit=unit, cut_pml=cut_pml)
    return gpu, ora


@pytest.mark.parametrize("unit", ["m/s", "s2/m2", "s2/km2"])
def test_param2wavenum_extension(unit):
    import torch
    from helm3d import Helm3d
    rng = np.random.default_rng(5)
    npml = [[2, 1, 3], [1, 2, 0]]
    n = int(np.prod(N_PHYS))
    v = rng.uniform(1500, 4000, n) if unit == "m/s" else rng.uniform(0.05, 0.4, n) * (1e-6 if unit == "s2/m2" else 1)
    H = Helm3d(device="cuda:0")
    p6 = (2, 1, 1, 2, 3, 0)
    wn, dwn, ddwn = H.param2wavenum(torch.from_numpy(v).cuda(), N_PHYS, p6, 4.5, unit, derivs=2)
    ext = F.extension(N_PHYS, npml) @ v
    ow, od, odd = O.param2wavenum(ext, 4.5, unit)
    assert relerr(wn.cpu().numpy(), ow) <= 1e-14
    assert relerr(dwn.cpu().numpy(), od) <= 1e-14
    if unit == "m/s":
        assert relerr(ddwn.cpu().numpy(), odd) <= 1e-14
    else:
        assert not ddwn.cpu().numpy().any()
    # Pext of an arbitrary field
    e = H.extend(torch.from_numpy(v).cuda(), N_PHYS, p6).cpu().numpy()
    assert np.array_equal(e, ext)


def test_extension_adjoint_kernel():
    import torch
    from helm3d import Helm3d
    rng = np.random.default_rng(6)
    for npml, p6 in (([[2, 1, 3], [1, 2, 0]], (2, 1, 1, 2, 3, 0)), ([[5, 5, 5], [5, 5, 5]], (5,) * 6)):
        P = F.extension(N_PHYS, npml)
        gp = rng.standard_normal(P.shape[0])
        H = Helm3d(device="cuda:0")
        g = H.extension_adjoint(torch.from_numpy(gp).cuda(), N_PHYS, p6).cpu().numpy()
        assert relerr(g, P.T @ gp) <= 1e-14
        g0 = torch.ones(P.shape[1], dtype=torch.float64, device="cuda:0")
        H.extension_adjoint(torch.from_numpy(gp).cuda(), N_PHYS, p6, g=g0, accumulate=True)
        assert relerr(g0.cpu().numpy(), P.T @ gp + 1.0) <= 1e-14


def test_sparse_rows_operator():
    import scipy.sparse as sp
    import torch
    from helm3d import Helm3d
    from helm3d.core import SparseRows
    rng = np.random.default_rng(8)
    A = sp.random(300, 500, density=0.01, random_state=9, format="csr")
    A = A + sp.csr_matrix((np.ones(3), ([0, 299, 150], [0, 499, 3])), shape=A.shape)
    op = SparseRows(A, "cuda:0")
    H = Helm3d(device="cuda:0")
    x = rng.standard_normal(500) + 1j * rng.standard_normal(500)
    y0 = rng.standard_normal(300) + 1j * rng.standard_normal(300)
    xt = torch.from_numpy(x).cuda()
    out = torch.from_numpy(y0).cuda()
    H.csr_apply(op, xt, out, alpha=-2.0, accumulate=True)
    assert relerr(out.cpu().numpy(), y0 - 2.0 * (A @ x)) <= 1e-14
    out = torch.zeros(300, dtype=torch.complex128, device="cuda:0")
    H.csr_apply(op, xt, out)
    assert relerr(out.cpu().numpy(), A @ x) <= 1e-14


@pytest.mark.parametrize("grid", ["phys", "pml"])
@pytest.mark.parametrize("const_dwn", [False, True])
def test_gradient_kernel(grid, const_dwn):
    import torch
    from helm3d import Helm3d
    rng = np.random.default_rng(12)
    n = (70, 9, 13)
    npml = np.array([[5, 2, 3], [4, 2, 3]])
    h = (12.5, 12.5, 12.5)
    N = int(np.prod(n))
    wn = rng.uniform(2e-4, 2e-3, N)
    dwn = np.full(N, 1.0) if const_dwn else rng.uniform(-1e-3, 1e-3, N)
    U = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    V = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    scale = 3.5
    H = Helm3d(n, h, npml, torch.from_numpy(wn).cuda())
    full = scale * np.real(np.conj(U) * O.hmvp("adj_deriv", dwn, h, n, npml, V))
    if grid == "phys":
        ref = F.padding(H.n_phys, npml).T @ full
    else:
        ref = full
    g0 = rng.standard_normal(ref.size)
    g = torch.from_numpy(g0.copy()).cuda()
    H.grad_accum(torch.from_numpy(U).cuda(), torch.from_numpy(V).cuda(), g, scale=scale,
                 dwn=None if const_dwn else torch.from_numpy(dwn).cuda(), grid=grid)
    assert relerr(g.cpu().numpy() - g0, ref) <= 1e-13


@pytest.mark.parametrize("unit,cut_pml", [("s2/m2", True), ("s2/m2", False), ("m/s", False)])
def test_objective_and_gradient_vs_oracle(unit, cut_pml):
    gpu, ora = _pair(unit=unit, cut_pml=cut_pml)
    m_true, m0 = _models(unit)
    Dobs = ora.forward(m_true)
    print(f"objective unit={unit} cut_pml={cut_pml}", flush=True)
    f, g = gpu.objective(m0, Dobs)
    fo, go = ora.objective(m0, Dobs)
    assert abs(f - fo) <= 1e-8 * abs(fo)
    assert relerr(g.cpu().numpy(), go) <= 1e-8


def test_pdefunc_modes_vs_oracle():
    import torch
    from helm3d.fwi import PDEfunc, FORW_MODEL, JACOB_FORW, JACOB_ADJ, HESS_GN
    gpu, ora = _pair(unit="s2/m2", cut_pml=False, freq=(5.0,))
    _, m0 = _models("s2/m2")
    rng = np.random.default_rng(13)
    dm = rng.standard_normal(m0.size) * m0.max()
    d = PDEfunc(FORW_MODEL, m0, None, None, None, gpu.model, problem=gpu).cpu().numpy()
    assert relerr(d, ora.forward(m0)) <= 1e-8
    jd = PDEfunc(JACOB_FORW, m0, None, dm, None, gpu.model, problem=gpu).cpu().numpy()
    assert relerr(jd, ora.jacob_forw(m0, dm)) <= 1e-8
    dd = rng.standard_normal(jd.shape) + 1j * rng.standard_normal(jd.shape)
    ja = PDEfunc(JACOB_ADJ, m0, None, dd, None, gpu.model, problem=gpu).cpu().numpy()
    assert relerr(ja, ora.jacob_adj(m0, dd)) <= 1e-8
    hg = PDEfunc(HESS_GN, m0, None, dm, None, gpu.model, problem=gpu).cpu().numpy()
    assert relerr(hg, ora.hess_gn(m0, dm)) <= 1e-8
    # adjoint identity on the device results
    assert abs(np.real(np.vdot(dd, jd)) - np.dot(dm, ja)) <= 1e-8 * abs(np.dot(dm, ja))


@pytest.mark.parametrize("unit,cut_pml", [("s2/m2", False), ("s2/m2", True), ("m/s", False)])
def test_hess_vs_oracle(unit, cut_pml):
    """PDEfunc(HESS) at a model with a non-zero data residual: the two
    second-order wavefields and the three gradient terms against the oracle."""
    from helm3d.fwi import PDEfunc, HESS
    gpu, ora = _pair(unit=unit, cut_pml=cut_pml, freq=(5.0,))
    m_true, m0 = _models(unit)
    Dobs = ora.forward(m_true)
    rng = np.random.default_rng(17)
    dm = rng.standard_normal(m0.size) * (m0.max() - m0.min())
    h = PDEfunc(HESS, m0, None, dm, Dobs, gpu.model, problem=gpu).cpu().numpy()
    ho = ora.hess(m0, dm, Dobs)
    assert relerr(h, ho) <= 1e-8
    # linear in the model perturbation (the reference's dV term uses the
    # forward derivative T(V) dm, so the product is not symmetric in dm)
    dm2 = rng.standard_normal(m0.size) * (m0.max() - m0.min())
    h2 = PDEfunc(HESS, m0, None, dm2, Dobs, gpu.model, problem=gpu).cpu().numpy()
    h12 = PDEfunc(HESS, m0, None, dm + 2.0 * dm2, Dobs, gpu.model, problem=gpu).cpu().numpy()
    assert relerr(h12, h + 2.0 * h2) <= 1e-8


def test_gradient_finite_differences_20_directions():
    """Central differences of the device objective on 20 seeded directions
    (solves at tol 1e-10), 24x22x20 physical grid + 5-point PML."""
    import torch
    from helm3d.fwi import FWIModel, FWIOpts, FWIProblem
    n_phys = (24, 22, 20)
    d = (25.0, 25.0, 25.0)
    model = FWIModel(o=(0.0, 0.0, 0.0), d=d, n=n_phys, freq=[6.0, 8.0],
                     xsrc=np.array([150.0, 400.0]), ysrc=np.array([250.0]), zsrc=np.array([50.0]),
                     xrec=25.0 * np.arange(2, 22, 3), yrec=25.0 * np.arange(2, 20, 4), zrec=np.array([75.0]),
                     f0=10.0, unit="s2/m2")
    P = FWIProblem(model, FWIOpts(pml=5, cut_pml=False, tol=1e-10, maxit=100000), device=torch.device("cuda:0"))
    m_true, m0 = _models("s2/m2", seed=21, n_phys=n_phys)
    Dobs = P.forw_model(m_true)
    f0, g = P.objective(m0, Dobs)
    g = g.cpu().numpy()
    rng = np.random.default_rng(20261207)
    worst = 0.0
    for it in range(20):
        dm = rng.standard_normal(m0.size) * m0.max()
        eps = 1e-4
        fp, _ = P.objective(m0 + eps * dm, Dobs, need_grad=False)
        fm, _ = P.objective(m0 - eps * dm, Dobs, need_grad=False)
        fd = (fp - fm) / (2 * eps)
        err = abs(fd - np.dot(g, dm)) / abs(fd)
        print(f"fd direction {it}: fd={fd:.9e} g.dm={np.dot(g, dm):.9e} rel={err:.2e}", flush=True)
        worst = max(worst, err)
    assert worst <= 1e-4


@pytest.mark.parametrize("wavefields", ["store", "recompute", "auto"])
def test_autograd_function(wavefields):
    import torch
    from helm3d.fwi import fwi_misfit
    gpu, ora = _pair(unit="s2/m2", cut_pml=True)
    m_true, m0 = _models("s2/m2")
    Dobs = ora.forward(m_true)
    f_ref, g_ref = gpu.objective(m0, Dobs)
    m = torch.from_numpy(m0.reshape(N_PHYS[::-1]).copy()).cuda().requires_grad_(True)
    f = fwi_misfit(m, gpu, Dobs, wavefields=wavefields)
    (2.0 * f).backward()
    assert abs(float(f) - f_ref) <= 1e-12 * abs(f_ref)
    assert m.grad.shape == m.shape
    assert relerr(m.grad.reshape(-1).cpu().numpy(), 2.0 * g_ref.cpu().numpy()) <= 1e-12


def test_mixed_precision_gradient():
    gpu, ora = _pair(unit="s2/m2", cut_pml=True, tol=1e-10, solver="mixed")
    m_true, m0 = _models("s2/m2")
    Dobs = ora.forward(m_true)
    f, g = gpu.objective(m0, Dobs)
    fo, go = ora.objective(m0, Dobs)
    assert abs(f - fo) <= 1e-8 * abs(fo)
    assert relerr(g.cpu().numpy(), go) <= 1e-6


@pytest.mark.parametrize("solver", ["fgmres", "mixed"])
def test_batched_objective_matches_unbatched(solver):
    """Sources of a frequency solved together (helm3d_batch_solve) give the
    objective and gradient of one-at-a-time solves."""
    import torch
    from helm3d.fwi import FWIModel, FWIOpts, FWIProblem, fwi_misfit
    model = FWIModel(o=(0.0, 0.0, 0.0), d=D, n=N_PHYS, freq=[5.0, 6.0], xsrc=np.array([100.0, 200.0, 300.0]),
                     ysrc=SRC[1], zsrc=SRC[2], xrec=REC[0], yrec=REC[1], zrec=REC[2], f0=10.0, unit="s2/m2")
    m_true, m0 = _models("s2/m2")
    P1 = FWIProblem(model, FWIOpts(pml=3, tol=1e-10, maxit=100000, solver=solver), device=torch.device("cuda:0"))
    Pb = FWIProblem(model, FWIOpts(pml=3, tol=1e-10, maxit=100000, solver=solver, batch=2, inner_iters=10),
                    device=torch.device("cuda:0"))
    Dobs = P1.forw_model(m_true)
    f1, g1 = P1.objective(m0, Dobs)
    fb, gb = Pb.objective(m0, Dobs)
    print(f"batched {solver}: f {f1} {fb}", flush=True)
    assert abs(f1 - fb) <= 1e-8 * abs(f1)
    assert relerr(gb.cpu().numpy(), g1.cpu().numpy()) <= 1e-7
    m = torch.from_numpy(m0.copy()).cuda().requires_grad_(True)
    fwi_misfit(m, Pb, Dobs, wavefields="recompute").backward()
    assert relerr(m.grad.cpu().numpy(), g1.cpu().numpy()) <= 1e-7


@pytest.mark.parametrize("unit,precision", [("s2/m2", "c128"), ("m/s", "c128"), ("s2/m2", "mixed")])
def test_multigrid_solver_objective(unit, precision):
    """PDEfunc(OBJ) with every solve preconditioned by the ML_GMRES V-cycle
    (FWIOpts.solver = "mg", hierarchy rebuilt per frequency, reloaded when the
    model changes) matches the oracle's exact solves."""
    gpu, ora = _pair(unit=unit, cut_pml=True, tol=1e-12, solver="mg")
    gpu.opts.mg_precision = precision
    m_true, m0 = _models(unit)
    Dobs = ora.forward(m_true)
    print(f"multigrid objective unit={unit}", flush=True)
    f, g = gpu.objective(m0, Dobs)
    fo, go = ora.objective(m0, Dobs)
    its = [i.iters for i in gpu.opts.stats]
    print("outer iterations per solve", min(its), max(its), flush=True)
    assert all(i.converged for i in gpu.opts.stats)
    assert abs(f - fo) <= 1e-8 * abs(fo)
    assert relerr(g.cpu().numpy(), go) <= 1e-8
    # second model on the same hierarchy
    f2, g2 = gpu.objective(m_true, Dobs)
    assert f2 <= 1e-16 * max(1.0, abs(fo)) + 1e-10 * abs(fo)
