"""CPU oracle for the 3D Helmholtz hot path (TEST INFRASTRUCTURE).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module.  It is the checker: nothing in the
product package imports, links or executes anything under ``oracle/``.

Contents
  * ctypes bindings of ``liboracle.so`` (C restatement of ``do_Hmvp`` for the
    six MEX variants, and of ``FGMRES.m``) and of ``_ref/libref_*.so`` (the
    reference stencil header compiled in place by ``oracle/Makefile``);
  * numpy restatements of the setup steps on the path:
      - ``param2wavenum``   tools/algorithms/CommonFreqModeling/param2wavenum.m:20-49
      - ``extend_model``    opExtension mode 1 (edge replication),
                            tools/operators/misc/opExtension.m:94-96, applied per
                            axis as in discrete_helmholtz.m:214-220,240
      - ``truncate``        comp2phys = Ppad' (cut_pml), discrete_helmholtz.m:231-233,
                            opExtension.m:89-90 (adjoint of zero padding)
      - ``fwi_wavelet``     tools/algorithms/CommonFreqModeling/fwi_wavelet.m:24-28
      - ``ls_misfit``       tools/algorithms/CommonFreqModeling/LS_misfit.m:21-26
      - ``kacz_sweep``      Kaczmarz double sweeps (Helm3dkaczswp_mex.c:60-117)
                            in the reference's lexicographic order and in the
                            27-colour order (C restatement in helm27_oracle.c)
      - ``cgmn``            CGMN.m with the Kaczmarz preconditioned system of
                            linearsolve.m:52-58
    (the PDEfunc restatement lives in fwi_oracle.py).

Parity status of each function is stated in its docstring; the C parts are
pinned against ``oracle/_ref`` (see ``tests/test_oracle_stencil.py``).
"""
from __future__ import annotations

import ctypes
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
VARIANTS = ("forw", "adj", "forw_deriv", "adj_deriv", "forw_wnc", "adj_wnc")

_dp = ctypes.POINTER(ctypes.c_double)
_HMVP_ARGS = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int]
HMVP_FN = ctypes.CFUNCTYPE(None, *_HMVP_ARGS)

_lib = None
_ref = {}


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def lib():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        for v in VARIANTS:
            f = getattr(L, "oracle_Hmvp_" + v)
            f.argtypes = _HMVP_ARGS
            f.restype = None
        L.oracle_pml_axis.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.oracle_hconst.argtypes = [_dp, _dp]
        L.oracle_fgmres.restype = ctypes.c_long
        L.oracle_fgmres.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, ctypes.c_int,
                                    ctypes.c_long, _dp, _dp, _dp, _dp, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_long, ctypes.c_long,
                                    _dp, ctypes.c_long, ctypes.POINTER(ctypes.c_long),
                                    _dp, ctypes.c_long, ctypes.POINTER(ctypes.c_long), _dp]
        for nm in ("oracle_kacz_sweep_lex", "oracle_kacz_sweep_mc"):
            f = getattr(L, nm)
            f.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, _dp, _dp, _dp, _dp]
            f.restype = None
        L.oracle_row.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 _dp, ctypes.POINTER(ctypes.c_long)]
        L.oracle_row.restype = ctypes.c_int
        _lib = L
    return _lib


def ref_available() -> bool:
    return all(os.path.exists(os.path.join(HERE, "_ref", f"libref_{v}.so")) for v in VARIANTS)


def ref_lib(variant: str):
    """The reference's do_Hmvp for one compile-time variant (oracle/_ref)."""
    if variant not in _ref:
        path = os.path.join(HERE, "_ref", f"libref_{variant}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: build with `make -C oracle ref` where the reference tree exists")
        L = ctypes.CDLL(path, mode=os.RTLD_LOCAL | getattr(os, "RTLD_NOW", 2))
        L.do_Hmvp.argtypes = _HMVP_ARGS
        L.do_Hmvp.restype = None
        _ref[variant] = L
    return _ref[variant]


def variant_name(mode: int, deriv: bool, wn_complex: bool) -> str:
    """MATLAB dispatch of helm3d_operto27_mvp.m:22-42."""
    base = "forw" if mode == 1 else "adj"
    if deriv:
        return base + "_deriv"
    return base + ("_wnc" if wn_complex else "")


def _prep(wn, h, n, npml, x):
    n = np.asarray(n, dtype=np.float64).reshape(3)
    N = int(np.prod(n))
    wn = np.asarray(wn).reshape(-1)
    wnr = np.ascontiguousarray(wn.real, dtype=np.float64)
    wni = np.ascontiguousarray(wn.imag if np.iscomplexobj(wn) else np.zeros_like(wnr), dtype=np.float64)
    x = np.asarray(x).reshape(-1)
    if x.size != N or wnr.size != N:
        raise ValueError("x and wn must have prod(n) elements")
    xr = np.ascontiguousarray(x.real, dtype=np.float64)
    xi = np.ascontiguousarray(x.imag if np.iscomplexobj(x) else np.zeros_like(xr), dtype=np.float64)
    h = np.ascontiguousarray(np.asarray(h, dtype=np.float64).reshape(3))
    # npml is 2x3 column-major [lo;hi] per axis -> flat [xlo,xhi,ylo,yhi,zlo,zhi]
    npml = np.asarray(npml, dtype=np.float64)
    if npml.shape == (2, 3):
        npml = npml.T.reshape(-1)
    npml = np.ascontiguousarray(npml.reshape(6))
    return wnr, wni, h, n, npml, xr, xi, N


def hmvp(variant: str, wn, h, n, npml, x, nthreads: int = 0, use_ref: bool = False):
    """y = H x (or variant) on split-complex arrays; returns complex128 (N,)."""
    wnr, wni, h, n, npml, xr, xi, N = _prep(wn, h, n, npml, x)
    yr = np.zeros(N)
    yi = np.zeros(N)
    nt = nthreads or os.cpu_count() or 1
    if use_ref:
        f = ref_lib(variant).do_Hmvp
    else:
        f = getattr(lib(), "oracle_Hmvp_" + variant)
    f(_p(wnr), _p(wni), _p(h), _p(n), _p(npml), _p(yr), _p(yi), _p(xr), _p(xi), int(nt))
    return yr + 1j * yi


def pml_axis(n: int, nlo: int, nhi: int):
    """xi_lo, xi_hi for 0-based indices -1..n (n+2 values each)."""
    lo = np.zeros(2 * (n + 2))
    hi = np.zeros(2 * (n + 2))
    lib().oracle_pml_axis(int(n), int(nlo), int(nhi), _p(lo), _p(hi))
    return lo[0::2] + 1j * lo[1::2], hi[0::2] + 1j * hi[1::2]


def hconst(h):
    out = np.zeros(8)
    lib().oracle_hconst(_p(np.ascontiguousarray(np.asarray(h, dtype=np.float64))), _p(out))
    keys = ("wn_coef", "cx", "cy", "cz", "xy", "xz", "yz", "w3a2")
    return dict(zip(keys, out))


def fgmres(variant: str, wn, h, n, npml, b, x0=None, tol=1e-6, maxinnerit=20, maxit=10000,
           nthreads: int = 0, use_ref: bool = False, max_iters_cap: int = 0):
    """Restarted GMRES of FGMRES.m (identity preconditioner) on the oracle or the
    reference stencil.  Returns (x, info) with info = dict(iters, relres, hst, res)."""
    wnr, wni, h, n, npml, br, bi, N = _prep(wn, h, n, npml, b)
    if x0 is None:
        xr = np.zeros(N)
        xi = np.zeros(N)
    else:
        x0 = np.asarray(x0).reshape(-1)
        xr = np.ascontiguousarray(x0.real, dtype=np.float64).copy()
        xi = np.ascontiguousarray(x0.imag, dtype=np.float64).copy()
    nt = nthreads or os.cpu_count() or 1
    if use_ref:
        fp = ctypes.cast(ref_lib(variant).do_Hmvp, ctypes.c_void_p)
    else:
        fp = ctypes.cast(getattr(lib(), "oracle_Hmvp_" + variant), ctypes.c_void_p)
    cap = 200000
    hst = np.zeros(cap)
    res = np.zeros(cap)
    nh = ctypes.c_long(0)
    nr = ctypes.c_long(0)
    rel = np.zeros(1)
    iters = lib().oracle_fgmres(fp, _p(wnr), _p(wni), _p(h), _p(n), _p(npml), int(nt), N,
                                _p(br), _p(bi), _p(xr), _p(xi), float(tol), int(maxinnerit),
                                int(maxit), int(max_iters_cap), _p(hst), cap, ctypes.byref(nh),
                                _p(res), cap, ctypes.byref(nr), _p(rel))
    info = dict(iters=int(iters), relres=float(rel[0]), hst=hst[:nh.value].copy(),
                res=res[:nr.value].copy())
    return xr + 1j * xi, info


# ---------------------------------------------------------------------------
# numpy restatements of the setup / gradient steps
# ---------------------------------------------------------------------------

def param2wavenum(v, freq, unit):
    """param2wavenum.m:20-49: returns (wn, dwn, ddwn)."""
    c = (2 * np.pi * freq) ** 2
    v = np.asarray(v, dtype=np.float64)
    if unit == "m/s":
        return c * v ** (-2), -2 * c * v ** (-3), 6 * c * v ** (-4)
    if unit == "s2/m2":
        return c * v, c * np.ones_like(v), np.zeros_like(v)
    if unit == "s2/km2":
        return 1e-6 * c * v, 1e-6 * c * np.ones_like(v), np.zeros_like(v)
    raise ValueError("Unknown unit - " + unit)


def extend_model(v, npml):
    """Edge-replicating extension into the PML (opExtension.m:94-96 per axis,
    composed as in discrete_helmholtz.m:214-217).  v is (nx,ny,nz) with x
    fastest in memory order; npml is 2x3 [lo;hi] per axis."""
    npml = np.asarray(npml).reshape(2, 3)
    v = np.asarray(v)
    pad = [(int(npml[0, a]), int(npml[1, a])) for a in range(3)]
    return np.pad(v, pad, mode="edge")


def truncate(u, npml):
    """comp2phys with cut_pml (Ppad', discrete_helmholtz.m:231-233)."""
    npml = np.asarray(npml).reshape(2, 3)
    nt = u.shape
    sl = tuple(slice(int(npml[0, a]), nt[a] - int(npml[1, a])) for a in range(3))
    return u[sl]


def fwi_wavelet(f, t0, f0):
    """fwi_wavelet.m:24-28 (Ricker spectrum with firing time t0)."""
    f = np.asarray(f, dtype=np.float64)
    w = np.exp(-1j * 2 * np.pi * f * t0)
    if f0:
        w = f ** 2 * np.exp(-(f / f0) ** 2) * w
    return w


def ls_misfit(x, y):
    """LS_misfit.m:21-26 (no offset mask): f = 1/2 ||x-y||^2, g = x - y."""
    r = x - y
    return 0.5 * float(np.vdot(r, r).real), r


def bicgstab(variant: str, wn, h, n, npml, b, tol=1e-6, maxit=10000, nthreads: int = 0):
    """BiCGStab of BICGSTAB.m:37-74 (identity preconditioner, x0 = 0) on the
    oracle stencil, numpy vector algebra.  beta = (rho_new/rho_old)(alpha/omega):
    the textbook update -- BICGSTAB.m:59 divides (rnew, r0_bar) by itself
    (beta = alpha/omega), a quirk this checker does not reproduce.  Returns
    (x, iters, relres_history) with iters counting iterations (2 matvecs each)."""
    A = lambda v: hmvp(variant, wn, h, n, npml, v, nthreads=nthreads)  # noqa: E731
    b = np.asarray(b, dtype=np.complex128).reshape(-1)
    x = np.zeros_like(b)
    r = b.copy()
    rh = r.copy()
    p = r.copy()
    nb = np.linalg.norm(b)
    rho = np.vdot(rh, r)
    hist = []
    it = 0
    while it < maxit:
        v = A(p)
        alpha = rho / np.vdot(rh, v)
        s = r - alpha * v
        t = A(s)
        omega = np.vdot(t, s) / np.vdot(t, t).real
        x = x + alpha * p + omega * s
        r = s - omega * t
        it += 1
        hist.append(np.linalg.norm(r) / nb)
        if hist[-1] < tol:
            break
        rho_new = np.vdot(rh, r)
        beta = (rho_new / rho) * (alpha / omega)
        rho = rho_new
        p = r + beta * (p - omega * v)
    return x, it, np.array(hist)


# ---------------------------------------------------------------------------
# Kaczmarz sweeps and CGMN
# ---------------------------------------------------------------------------

_KVAR = {"forw": 0, "adj": 1, "forw_wnc": 4, "adj_wnc": 5}


def kacz_sweep(variant: str, wn, h, n, npml, omega, x, b, order: str = "mc"):
    """One double Kaczmarz sweep (forward then backward) of x for A x = b with
    A = H (forw*) or H^H (adj*): Helm3dkaczswp_mex.c:60-117 in the reference's
    lexicographic order (order="lex") or the 27-colour order ("mc").
    Returns the swept copy of x."""
    wnr, wni, h, n, npml, xr, xi, N = _prep(wn, h, n, npml, x)
    xr = xr.copy()
    xi = (xi.copy() if xi is not None else np.zeros(N))
    b = np.asarray(b).reshape(-1)
    br = np.ascontiguousarray(b.real, dtype=np.float64)
    bi = np.ascontiguousarray(b.imag, dtype=np.float64)
    if wni is None:
        wni = np.zeros(N)
    f = lib().oracle_kacz_sweep_mc if order == "mc" else lib().oracle_kacz_sweep_lex
    f(_KVAR[variant], _p(wnr), _p(wni), _p(h), _p(n), _p(npml), float(omega), _p(xr), _p(xi), _p(br), _p(bi))
    return xr + 1j * xi


def row(variant: str, wn, h, n, npml, i, j, k):
    """Row (i,j,k) of H or H^H: (coefficients, flat neighbour indices)."""
    wnr, wni, h, n, npml, _, _, N = _prep(wn, h, n, npml, np.zeros(int(np.prod(n))))
    if wni is None:
        wni = np.zeros(N)
    c = np.zeros(54)
    idx = (ctypes.c_long * 27)()
    cnt = lib().oracle_row(_KVAR[variant], _p(wnr), _p(wni), _p(h), _p(n), _p(npml), int(i), int(j), int(k),
                           _p(c), idx)
    return c[:2 * cnt:2] + 1j * c[1:2 * cnt:2], np.array(idx[:cnt])


def cgmn(sweep, b, tol=1e-6, maxit=10000):
    """CGMN.m on the Kaczmarz-preconditioned system of linearsolve.m:52-58:
    b' = SWP(0, b), A' y = y - SWP(y, 0), CG from x = 0.  sweep(x, rhs) is one
    double sweep.  Returns (x, iterations, relative residual history)."""
    zb = np.zeros_like(b)
    bp = sweep(np.zeros_like(b), b)
    A = lambda y: y - sweep(y, zb)  # noqa: E731
    x = np.zeros_like(b)
    r = bp.copy()
    p = r.copy()
    normb = np.linalg.norm(bp)
    normr = np.linalg.norm(r)
    hst = [normr / normb]
    k = 0
    while True:
        Ap = A(p)
        alpha = normr ** 2 / np.vdot(p, Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        normr_old, normr = normr, np.linalg.norm(r)
        hst.append(normr / normb)
        k += 1
        if k >= maxit or hst[-1] <= tol:
            break
        p = r + (normr / normr_old) ** 2 * p
    return x, k, np.array(hst)


def crmn(sweep, b, tol=1e-6, maxit=10000):
    """CRMN.m:46-108 (conjugate residuals) on the Kaczmarz-preconditioned
    system of linearsolve.m:52-58 (b' = SWP(0, b), A' y = y - SWP(y, 0)),
    from x = 0.  Returns (x, iterations, relative residual history)."""
    zb = np.zeros_like(b)
    bp = sweep(np.zeros_like(b), b)
    A = lambda y: y - sweep(y, zb)  # noqa: E731
    x = np.zeros_like(b)
    r = bp.copy()
    Ar = A(r)
    p, Ap = r.copy(), Ar.copy()
    nb = np.linalg.norm(bp)
    rAr = np.vdot(r, Ar)
    nAp = np.linalg.norm(Ap)
    hst = [np.linalg.norm(r) / nb]
    k = 0
    while True:
        gamma = rAr / nAp ** 2
        x = x + gamma * p
        r = r - gamma * Ap
        hst.append(np.linalg.norm(r) / nb)
        k += 1
        if k >= maxit or hst[-1] <= tol:
            break
        Ar = A(r)
        old, rAr = rAr, np.vdot(r, Ar)
        beta = rAr / old
        p = r + beta * p
        Ap = Ar + beta * Ap
        nAp = np.linalg.norm(Ap)
    return x, k, np.array(hst)


# ---------------------------------------------------------------------------
# multigrid grid transfers
# ---------------------------------------------------------------------------

def lin_interp_1d(xin, xout):
    """LinInterp1D.m: weights (len(xout) x len(xin)) of linear interpolation,
    the bracket found from the nearest input point (stepped left when that
    point lies right of the output point)."""
    import scipy.sparse as sp
    xin = np.sort(np.asarray(xin, dtype=np.float64).reshape(-1))
    xout = np.sort(np.asarray(xout, dtype=np.float64).reshape(-1))
    nin = xin.size
    rows, cols, vals = [], [], []
    for r, x in enumerate(xout):
        I = int(np.argmin(np.abs(xin - x)))
        if xin[I] > x:
            I -= 1
        x1 = xin[I + 1] if I + 1 < nin else 1e12
        a = (x1 - x) / (x1 - xin[I])
        b = (x - xin[I]) / (x1 - xin[I])
        rows += [r, r]
        cols += [I, min(I + 1, nin - 1)]
        vals += [a, b]
    return sp.csr_matrix((vals, (rows, cols)), shape=(xout.size, nin))


def prolongation(nf, nc):
    """c2f of fine2coarse(nf, nc, 'linear') (fine2coarse.m:49-53): the
    trilinear interpolation from the coarse to the fine grid, x fastest."""
    import scipy.sparse as sp
    L = [lin_interp_1d(np.linspace(0, 1, nc[a]), np.linspace(0, 1, nf[a])) for a in range(3)]
    return sp.kron(L[2], sp.kron(L[1], L[0])).tocsr()
