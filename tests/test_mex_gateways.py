"""The MEX gateways of slim-release-apps_amd/mex/, compiled against a test
stand-in of MATLAB's mx / mex API (tests/mexshim) and driven through
mexFunction as MATLAB would call them.

CPU (no device needed): every gateway builds and loads; the argument checks
raise the reference's error identifiers (Helm3dmvp_mex.c:75-85,
Helm3dkaczswp_mex.c:148-165 -- Nsize, Xsize, Bsize) and the added ones
(nargin, WNsize, Hsize, NPMLsize, b columns); an error leaves no gateway
allocation behind for MATLAB to collect.

GPU: the six Helm3dmvp_*_mex gateways and the six exact-signature
helm3d_do_Hmvp_* entry points against the reference stencil compiled in place
(oracle/_ref; the oracle restatement where _ref is absent) at <= 1e-12, with
complex and real x; the three sweep gateways on P = 3 columns against the
oracle's 27-colour sweep per column; the solve gateway (one column: FGMRES.m,
three columns: FGMRES_MT.m) with every residual checked by the oracle
operator, the context released by its exit handler.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle as O

BUILD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mexshim", "_build")
ID = "SLIM_release_apps:tools:algorithms:ThreeDFreqModeling:"
MVP = ("forw", "adj", "forw_deriv", "adj_deriv", "forw_wnc", "adj_wnc")
SWEEPS = {"Helm3d_kacz_sweep_mex": "forw", "Helm3d_kacz_sweep_adj_mex": "adj",
          "Helm3d_kacz_sweep_wnc_mex": "forw_wnc"}
NAMES = [f"Helm3dmvp_{v}_mex" for v in MVP] + list(SWEEPS) + ["Helm3d_solve_gpu_mex"]


class MexError(Exception):
    def __init__(self, ident, msg):
        super().__init__(f"{ident}: {msg}")
        self.identifier = ident


class Gateway:
    """One compiled MEX file and the harness of mexshim.c."""

    def __init__(self, name):
        import helm3d
        helm3d.lib()  # the product library (and torch's HIP runtime) first
        path = os.path.join(BUILD, name + ".so")
        if not os.path.exists(path):
            pytest.fail(f"{path} is not built (make -C tests/mexshim)")
        L = ctypes.CDLL(path)
        vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
        for fn, res, args in (("shim_matrix", vp, [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]),
                              ("shim_pr", dp, [vp]), ("shim_pi", dp, [vp]), ("shim_m", ctypes.c_size_t, [vp]),
                              ("shim_n", ctypes.c_size_t, [vp]), ("shim_free", None, [vp]),
                              ("shim_err_id", ctypes.c_char_p, []), ("shim_err_msg", ctypes.c_char_p, []),
                              ("shim_auto_freed", ctypes.c_int, []), ("shim_arrays_dropped", ctypes.c_int, []),
                              ("shim_call", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp), ctypes.c_int,
                                                           ctypes.POINTER(vp)]),
                              ("shim_exit", ctypes.c_int, [])):
            f = getattr(L, fn)
            f.restype = res
            f.argtypes = args
        self.L = L

    def _mat(self, a):
        a = np.asarray(a)
        if a.ndim == 0:
            a = a.reshape(1, 1)
        elif a.ndim == 1:
            a = a.reshape(-1, 1)
        m, n = a.shape
        cplx = np.iscomplexobj(a)
        p = self.L.shim_matrix(m, n, int(cplx))
        if m * n:
            np.ctypeslib.as_array(self.L.shim_pr(p), shape=(m * n,))[:] = a.real.reshape(-1, order="F")
            if cplx:
                np.ctypeslib.as_array(self.L.shim_pi(p), shape=(m * n,))[:] = a.imag.reshape(-1, order="F")
        return p

    def _out(self, p):
        m, n = self.L.shim_m(p), self.L.shim_n(p)
        re = np.ctypeslib.as_array(self.L.shim_pr(p), shape=(m * n,)).copy() if m * n else np.zeros(0)
        pi = self.L.shim_pi(p)
        out = re + 1j * np.ctypeslib.as_array(pi, shape=(m * n,)) if pi and m * n else re
        self.L.shim_free(p)
        return out.reshape((m, n), order="F")

    def __call__(self, *args, nlhs=1):
        ins = [self._mat(a) for a in args]
        prhs = (ctypes.c_void_p * max(1, len(ins)))(*ins)
        plhs = (ctypes.c_void_p * max(1, nlhs))()
        raised = self.L.shim_call(nlhs, plhs, len(ins), prhs)
        for p in ins:
            self.L.shim_free(p)
        if raised:
            raise MexError(self.L.shim_err_id().decode(), self.L.shim_err_msg().decode())
        return [self._out(plhs[i]) for i in range(max(1, nlhs)) if plhs[i]]

    def exit(self):
        return self.L.shim_exit()


def _case(n=(9, 8, 7), npml=((2, 2, 1), (3, 2, 2)), seed=0):
    rng = np.random.default_rng(seed)
    N = int(np.prod(n))
    wn = rng.uniform(2e-4, 2e-3, N)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    return np.array(n, dtype=float), np.array(npml, dtype=float), np.array([10.0, 11.0, 9.0]), wn, x


def _relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ------------------------------------------------------------------------ CPU

@pytest.mark.parametrize("name", NAMES)
def test_gateway_builds_and_loads(name):
    G = Gateway(name)
    assert G.L.mexFunction is not None


@pytest.mark.parametrize("variant", MVP)
def test_matvec_gateway_argument_errors(variant):
    G = Gateway(f"Helm3dmvp_{variant}_mex")
    n, npml, h, wn, x = _case()
    bad = [
        ((wn, h, n, npml, x), "Helm3dmvp_mex:nargin"),
        ((wn, h, np.array([[9.0, 8.0], [7.0, 1.0]]), npml, x, 1), "Helm3dmvp_mex:Nsize"),
        ((wn, h, np.array([9.0, 8.0]), npml, x, 1), "Helm3dmvp_mex:Nsize"),
        ((wn, h, np.array([9.0, 8.5, 7.0]), npml, x, 1), "Helm3dmvp_mex:Nsize"),
        ((wn, h, np.array([9.0, 0.0, 7.0]), npml, x, 1), "Helm3dmvp_mex:Nsize"),
        ((wn, h, n, npml, x[:-1], 1), "Helm3dkaczswp_mex:Xsize"),  # the reference's id (sic)
        ((wn[:-1], h, n, npml, x, 1), "Helm3dmvp_mex:WNsize"),
        ((wn, h[:2], n, npml, x, 1), "Helm3dmvp_mex:Hsize"),
        ((wn, h, n, npml[:, :2], x, 1), "Helm3dmvp_mex:NPMLsize"),
    ]
    for args, ident in bad:
        with pytest.raises(MexError) as e:
            G(*args)
        assert e.value.identifier == ID + ident, (args[2], e.value)
        assert G.L.shim_auto_freed() == 0
    # n given as a row or a column passes the shape check (Helm3dmvp_mex.c:75)
    for nn in (n.reshape(1, 3), n.reshape(3, 1)):
        with pytest.raises(MexError) as e:
            G(wn, h, nn, npml, x[:-1], 1)
        assert e.value.identifier.endswith("Xsize")


@pytest.mark.parametrize("name", list(SWEEPS))
def test_sweep_gateway_argument_errors(name):
    G = Gateway(name)
    n, npml, h, wn, x = _case()
    X = np.stack([x, 2 * x], axis=1)
    bad = [
        ((wn, h, n, npml, 1.5, X), "nargin"),
        ((wn, h, np.array([9.0, 8.0]), npml, 1.5, X, X), "Nsize"),
        ((wn, h, n, npml, 1.5, X[:-1], X), "Xsize"),
        ((wn, h, n, npml, 1.5, X, X[:-1]), "Bsize"),
        ((wn, h, n, npml, 1.5, X, X[:, :1]), "Bsize"),  # b columns must match x columns
        ((wn[:-1], h, n, npml, 1.5, X, X), "WNsize"),
    ]
    for args, ident in bad:
        with pytest.raises(MexError) as e:
            G(*args)
        assert e.value.identifier == ID + "Helm3dkaczswp_mex:" + ident, e.value
        assert G.L.shim_auto_freed() == 0


def test_solve_gateway_argument_errors():
    G = Gateway("Helm3d_solve_gpu_mex")
    n, npml, h, wn, x = _case()
    for args, ident in (((wn, h, n, npml), "nargin"),
                        ((wn, h, np.array([9.0, 8.0]), npml, x), "Nsize"),
                        ((wn, h, n, npml, x[:-1]), "Bsize"),
                        ((wn[:-1], h, n, npml, x), "WNsize")):
        with pytest.raises(MexError) as e:
            G(*args)
        assert e.value.identifier == ID + "Helm3d_solve_gpu_mex:" + ident, e.value
    assert G.exit() == 0  # no context was created


# ------------------------------------------------------------------------ GPU

def _ref(variant, wn, h, n, npml, x):
    return O.hmvp(variant, wn, h, tuple(int(a) for a in n), npml, x, use_ref=O.ref_available())


@pytest.mark.gpu
@pytest.mark.parametrize("variant", MVP)
@pytest.mark.parametrize("case", [0, 1])
def test_matvec_gateways_match_reference(variant, case):
    shape = [((9, 8, 7), ((2, 2, 1), (3, 2, 2))), ((12, 10, 9), ((0, 3, 2), (3, 2, 3)))][case]
    n, npml, h, wn, x = _case(*shape, seed=case)
    if variant.endswith("wnc"):
        wn = wn + 1j * np.random.default_rng(9).uniform(-2e-4, 2e-4, wn.size)
    G = Gateway(f"Helm3dmvp_{variant}_mex")
    (y,) = G(wn, h, n, npml, x, 4)
    assert y.shape == (x.size, 1) and np.iscomplexobj(y)
    assert _relerr(y[:, 0], _ref(variant, wn, h, n, npml, x)) <= 1e-12
    (yr,) = G(wn, h, n, npml, x.real, 4)  # a real x has zero imaginary part
    assert _relerr(yr[:, 0], _ref(variant, wn, h, n, npml, x.real + 0j)) <= 1e-12
    # only the first column of x is used (Helm3dmvp_mex.c:86-93)
    (y2,) = G(wn, h, n, npml, np.stack([x, 3 * x], axis=1), 4)
    assert np.array_equal(y2, y)
    assert G.L.shim_auto_freed() == 0 and G.L.shim_arrays_dropped() == 0


@pytest.mark.gpu
@pytest.mark.parametrize("variant", MVP)
def test_exact_signature_entry_points_match_reference(variant):
    import helm3d
    n, npml, h, wn, x = _case((10, 9, 8), ((3, 2, 2), (2, 3, 1)), seed=5)
    wni = np.random.default_rng(6).uniform(-2e-4, 2e-4, wn.size) if variant.endswith("wnc") else np.zeros(wn.size)
    P = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (wn, wni, h, n, npml.T.reshape(-1), x.real, x.imag)]
    yr, yi = np.zeros(wn.size), np.zeros(wn.size)
    f = getattr(helm3d.lib(), f"helm3d_do_Hmvp_{variant}")
    code = f(P(arrs[0]), P(arrs[1]), P(arrs[2]), P(arrs[3]), P(arrs[4]), P(yr), P(yi), P(arrs[5]), P(arrs[6]), 4)
    assert code == 0
    w = wn + 1j * wni if variant.endswith("wnc") else wn
    assert _relerr(yr + 1j * yi, _ref(variant, w, h, n, npml, x)) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(SWEEPS))
def test_sweep_gateways_on_columns(name):
    variant = SWEEPS[name]
    n, npml, h, wn, x = _case((10, 9, 8), ((2, 2, 2), (3, 2, 1)), seed=11)
    if variant.endswith("wnc"):
        wn = wn + 1j * np.random.default_rng(12).uniform(-2e-4, 2e-4, wn.size)
    rng = np.random.default_rng(13)
    X = np.stack([x, rng.standard_normal(x.size) + 0j, 1j * x], axis=1)
    B = rng.standard_normal((x.size, 3)) + 1j * rng.standard_normal((x.size, 3))
    G = Gateway(name)
    (Y,) = G(wn, h, n, npml, 1.3, X, B)
    assert Y.shape == (x.size, 3)
    nn = tuple(int(a) for a in n)
    for s in range(3):
        ys = O.kacz_sweep(variant, wn, h, nn, npml, 1.3, X[:, s], B[:, s])
        assert _relerr(Y[:, s], ys) <= 1e-12, s
    # real x and b: zero imaginary parts
    (Yr,) = G(wn, h, n, npml, 1.3, X.real, B.real)
    ys = O.kacz_sweep(variant, wn, h, nn, npml, 1.3, X[:, 0].real + 0j, B[:, 0].real + 0j)
    assert _relerr(Yr[:, 0], ys) <= 1e-12
    assert G.L.shim_auto_freed() == 0 and G.L.shim_arrays_dropped() == 0


@pytest.mark.gpu
def test_solve_gateway_one_and_many_columns():
    from helm3d.models import config_small, extend, point_source, vec, wavenumber
    c = config_small(n=(20, 18, 16), npml=5)
    wn = vec(wavenumber(extend(c.v, c.npml), c.freq))
    n = np.array(c.nt, dtype=float)
    npml = np.array(c.npml6, dtype=float).reshape(3, 2).T
    h = np.full(3, c.h)
    b = point_source(c, c.sources[0])
    G = Gateway("Helm3d_solve_gpu_mex")
    x, info = G(wn, h, n, npml, b, 1, 1e-6, 20, 10000, nlhs=2)
    assert x.shape == (b.size, 1) and info.shape == (4, 1)
    xo, io = O.fgmres("forw", wn, tuple(h), c.nt, npml, b, tol=1e-6, maxinnerit=20)
    r = b - O.hmvp("forw", wn, tuple(h), c.nt, npml, x[:, 0])
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    assert info[3, 0] == 1.0 and rel < 1e-6 and abs(rel - info[2, 0]) < 1e-9
    assert abs(info[0, 0] - io["iters"]) <= 0.02 * io["iters"] + 2
    # three columns, adjoint system, LinSolveOpts defaults (restart 5): FGMRES_MT
    rng = np.random.default_rng(3)
    Bm = np.stack([b, rng.standard_normal(b.size) + 1j * rng.standard_normal(b.size), 2 * b], axis=1)
    Xm, infm = G(wn, h, n, npml, Bm, 2, nlhs=2)
    assert Xm.shape == Bm.shape and infm.shape == (4, 3)
    for s in range(3):
        r = Bm[:, s] - O.hmvp("adj", wn, tuple(h), c.nt, npml, Xm[:, s])
        assert np.linalg.norm(r) / np.linalg.norm(Bm[:, s]) < 1e-6, s
    assert np.all(infm[0] == infm[0, 0])  # lockstep (FGMRES_MT.m)
    assert G.L.shim_auto_freed() == 0 and G.L.shim_arrays_dropped() == 0
    assert G.exit() == 1  # the exit handler releases the device context
