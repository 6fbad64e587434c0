"""Device-resident Helmholtz operator over the C ABI (torch tensors as buffers).

``Helm3d`` owns one ``helm3d_ctx``: the model (wn on the PML grid, PML
tables), the Krylov workspace, and the HIP stream it runs on (torch's current
stream at call time).  Vectors are interleaved complex torch tensors on the
context's device, x fastest (flattened in MATLAB column-major order).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import check, lib


def _stream_ptr() -> int:
    return torch.cuda.current_stream().cuda_stream


def _dptr(t: torch.Tensor) -> int:
    return t.data_ptr()


def _as_int3(v) -> tuple:
    v = tuple(int(a) for a in np.asarray(v).reshape(-1))
    if len(v) != 3:
        raise ValueError("expected 3 values")
    return v


def npml_flat(npml) -> tuple:
    """Accept the MATLAB 2x3 [lo;hi] matrix, a 6-vector [xlo,xhi,ylo,yhi,zlo,zhi]
    or a scalar (same count on every face)."""
    a = np.asarray(npml)
    if a.size == 1:
        return (int(a),) * 6
    if a.shape == (2, 3):
        return tuple(int(v) for v in a.T.reshape(-1))
    if a.size == 6:
        return tuple(int(v) for v in a.reshape(-1))
    raise ValueError("npml must be a scalar, a 2x3 [lo;hi] matrix or 6 values")


@dataclass
class SolveInfo:
    """Outcome of a solve (FGMRES.m outputs plus counters)."""
    iters: int
    cycles: int
    relres: float
    quasi_relres: float
    converged: bool
    inner_iters: int = 0
    hst: Optional[np.ndarray] = field(default=None, repr=False)


class Helm3d:
    """H(wn) on one device; forward / adjoint / derivative products and solves."""

    def __init__(self, n=None, h=None, npml=None, wn: Optional[torch.Tensor] = None, device=None):
        if device is None:
            device = wn.device if wn is not None else torch.device("cuda", torch.cuda.current_device())
        device = torch.device(device)
        if device.type != "cuda":
            raise RuntimeError("Helm3d runs on a HIP device; got " + str(device))
        self.device = device
        self._ctx = ctypes.c_void_p()
        with torch.cuda.device(device):
            check(lib().helm3d_create(device.index if device.index is not None else -1,
                                      ctypes.c_void_p(_stream_ptr()), ctypes.byref(self._ctx)))
        self.n = None
        if wn is not None:
            self.set_model(n, h, npml, wn)

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None) and self._ctx.value:
            lib().helm3d_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _bind(self):
        check(lib().helm3d_set_stream(self._ctx, ctypes.c_void_p(_stream_ptr())))

    # -- model --------------------------------------------------------------
    def set_model(self, n, h, npml, wn: torch.Tensor):
        self.n = _as_int3(n)
        self.h = tuple(float(a) for a in np.asarray(h, dtype=np.float64).reshape(3))
        self.npml = npml_flat(npml)
        self.N = int(np.prod(self.n))
        wn = wn.reshape(-1)
        if wn.numel() != self.N:
            raise ValueError("wn must have prod(n) elements")
        self.wn_complex = wn.is_complex()
        wn = wn.to(device=self.device, dtype=torch.complex128 if self.wn_complex else torch.float64).contiguous()
        n3 = (ctypes.c_int * 3)(*self.n)
        h3 = (ctypes.c_double * 3)(*self.h)
        p6 = (ctypes.c_int * 6)(*self.npml)
        with torch.cuda.device(self.device):
            self._bind()
            check(lib().helm3d_set_model(self._ctx, n3, h3, p6, ctypes.c_void_p(_dptr(wn)),
                                         int(self.wn_complex)))
        self.wn = wn

    @property
    def n_phys(self):
        p = self.npml
        return tuple(self.n[a] - p[2 * a] - p[2 * a + 1] for a in range(3))

    # -- products -----------------------------------------------------------
    def _vec(self, x: torch.Tensor, dtype) -> torch.Tensor:
        if x.device != self.device:
            raise ValueError(f"vector on {x.device}, operator on {self.device}")
        x = x.reshape(-1)
        if x.numel() != self.N:
            raise ValueError("vector must have prod(n) elements")
        return x.to(dtype).contiguous()

    def apply(self, x: torch.Tensor, adjoint: bool = False, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """y = H x (or H^H x).  complex128 input runs the FP64 kernel, complex64
        the FP32 kernel."""
        prec = _lib.HELM3D_C64 if x.dtype == torch.complex64 else _lib.HELM3D_C128
        dt = torch.complex64 if prec == _lib.HELM3D_C64 else torch.complex128
        x = self._vec(x, dt)
        y = out if out is not None else torch.empty_like(x)
        with torch.cuda.device(self.device):
            self._bind()
            check(lib().helm3d_apply(self._ctx, _lib.HELM3D_ADJ if adjoint else _lib.HELM3D_FORW, prec,
                                     ctypes.c_void_p(_dptr(x)), ctypes.c_void_p(_dptr(y))))
        return y

    def apply_deriv(self, dwn: torch.Tensor, x: torch.Tensor, adjoint: bool = False) -> torch.Tensor:
        """(dH . dwn) x or its adjoint (the DERIV stencil with real dwn)."""
        prec = _lib.HELM3D_C64 if x.dtype == torch.complex64 else _lib.HELM3D_C128
        dt = torch.complex64 if prec == _lib.HELM3D_C64 else torch.complex128
        rt = torch.float32 if prec == _lib.HELM3D_C64 else torch.float64
        x = self._vec(x, dt)
        dwn = dwn.reshape(-1).to(device=self.device, dtype=rt).contiguous()
        if dwn.numel() != self.N:
            raise ValueError("dwn must have prod(n) elements")
        y = torch.empty_like(x)
        with torch.cuda.device(self.device):
            self._bind()
            check(lib().helm3d_apply_deriv(self._ctx, int(adjoint), prec, ctypes.c_void_p(_dptr(dwn)),
                                           ctypes.c_void_p(_dptr(x)), ctypes.c_void_p(_dptr(y))))
        return y

    # -- solves -------------------------------------------------------------
    @staticmethod
    def _opts(tol, maxinnerit, maxit, use_x0, inner_tol=1e-2, inner_m=20, inner_maxit=200, inner="bicgstab"):
        o = _lib.SolveOpts()
        lib().helm3d_solve_opts_default(ctypes.byref(o))
        o.tol = float(tol)
        o.maxinnerit = int(maxinnerit)
        o.maxit = int(maxit)
        o.use_x0 = int(use_x0)
        o.inner_tol = float(inner_tol)
        o.inner_m = int(inner_m)
        o.inner_maxit = int(inner_maxit)
        o.inner_solver = {"bicgstab": _lib.HELM3D_INNER_BICGSTAB, "gmres": _lib.HELM3D_INNER_GMRES}[inner]
        return o

    def solve(self, b: torch.Tensor, mode: int = 1, tol: float = 1e-6, maxinnerit: int = 20,
              maxit: int = 10000, x0: Optional[torch.Tensor] = None, return_hst: bool = False,
              raise_on_notconv: bool = False):
        """FGMRES.m restarted GMRES (identity preconditioner) in complex128.
        mode 1: H x = b; otherwise H^H x = b (linearsolve.m:44-48)."""
        b = self._vec(b, torch.complex128)
        x = torch.zeros_like(b) if x0 is None else self._vec(x0, torch.complex128).clone()
        o = self._opts(tol, maxinnerit, maxit, x0 is not None)
        info = _lib.SolveInfo()
        cap = 0
        hst = None
        nh = ctypes.c_longlong(0)
        if return_hst:
            cap = int(min(maxit * maxinnerit + 1, 1 << 22))
            hst = np.zeros(cap)
        with torch.cuda.device(self.device):
            self._bind()
            code = lib().helm3d_solve(self._ctx, int(mode), ctypes.c_void_p(_dptr(b)), ctypes.c_void_p(_dptr(x)),
                                      ctypes.byref(o), ctypes.byref(info),
                                      hst.ctypes.data_as(_lib._dp) if hst is not None else None, cap,
                                      ctypes.byref(nh))
        check(code, allow=() if raise_on_notconv else (_lib.HELM3D_ERR_NOTCONV,))
        si = SolveInfo(int(info.iters), int(info.cycles), float(info.relres), float(info.quasi_relres),
                       bool(info.converged), hst=hst[:nh.value].copy() if hst is not None else None)
        return x, si

    def solve_mixed(self, b: torch.Tensor, mode: int = 1, tol: float = 1e-6, maxinnerit: int = 20,
                    maxit: int = 10000, inner_tol: float = 1e-2, inner_m: int = 20, inner_maxit: int = 200,
                    x0: Optional[torch.Tensor] = None, raise_on_notconv: bool = False, inner: str = "bicgstab"):
        """Mixed precision: FP64 flexible GMRES (FGMRES.m with a preconditioner)
        whose preconditioner is a complex64 solve of the same operator to
        inner_tol -- BiCGStab (default, linearsolve.m:76-77) or GMRES(inner_m)
        (linearsolve.m:81-82).  inner_iters counts complex64 matvecs."""
        b = self._vec(b, torch.complex128)
        x = torch.zeros_like(b) if x0 is None else self._vec(x0, torch.complex128).clone()
        o = self._opts(tol, maxinnerit, maxit, x0 is not None, inner_tol, inner_m, inner_maxit, inner)
        info = _lib.SolveInfo()
        with torch.cuda.device(self.device):
            self._bind()
            code = lib().helm3d_solve_mixed(self._ctx, int(mode), ctypes.c_void_p(_dptr(b)),
                                            ctypes.c_void_p(_dptr(x)), ctypes.byref(o), ctypes.byref(info))
        check(code, allow=() if raise_on_notconv else (_lib.HELM3D_ERR_NOTCONV,))
        return x, SolveInfo(int(info.iters), int(info.cycles), float(info.relres), float(info.quasi_relres),
                            bool(info.converged), inner_iters=int(info.inner_iters))

    def solve_bicgstab(self, b: torch.Tensor, mode: int = 1, tol: float = 1e-6, maxit: int = 10000,
                       raise_on_notconv: bool = False):
        """BiCGStab (BICGSTAB.m, LinSolveOpts.SOLVE_BICGSTAB) in complex128 from
        x = 0; info.iters counts matvecs, info.relres is the true FP64 residual."""
        b = self._vec(b, torch.complex128)
        x = torch.zeros_like(b)
        o = self._opts(tol, 20, maxit, False)
        info = _lib.SolveInfo()
        with torch.cuda.device(self.device):
            self._bind()
            code = lib().helm3d_solve_bicgstab(self._ctx, int(mode), ctypes.c_void_p(_dptr(b)),
                                               ctypes.c_void_p(_dptr(x)), ctypes.byref(o), ctypes.byref(info))
        check(code, allow=() if raise_on_notconv else (_lib.HELM3D_ERR_NOTCONV,))
        return x, SolveInfo(int(info.iters), int(info.cycles), float(info.relres), float(info.quasi_relres),
                            bool(info.converged))

    def kacz_sweep(self, x: torch.Tensor, b: Optional[torch.Tensor] = None, mode: int = 1, omega: float = 1.5,
                   nsweeps: int = 1) -> torch.Tensor:
        """nsweeps double Kaczmarz sweeps (27-colour order) of a copy of x for
        A x = b, A = H (mode 1) or H^H (Helm3dkaczswp_mex.c:60-117)."""
        y = self._vec(x, torch.complex128).clone()
        bb = self._vec(b, torch.complex128) if b is not None else None
        with torch.cuda.device(self.device):
            self._bind()
            check(lib().helm3d_kacz_sweep(self._ctx, int(mode), float(omega), int(nsweeps),
                                          ctypes.c_void_p(_dptr(y)),
                                          ctypes.c_void_p(_dptr(bb)) if bb is not None else None))
        return y

    def solve_cgmn(self, b: torch.Tensor, mode: int = 1, tol: float = 1e-6, maxit: int = 10000,
                   raise_on_notconv: bool = False, method: str = "cgmn"):
        """CGMN.m (or, with method "crmn", CRMN.m) on the Kaczmarz system
        (linearsolve.m:52-63, omega 1.5): iters counts CG / CR iterations,
        quasi_relres is the preconditioned residual, relres the true one."""
        fn = {"cgmn": lib().helm3d_solve_cgmn, "crmn": lib().helm3d_solve_crmn}[method]
        b = self._vec(b, torch.complex128)
        x = torch.zeros_like(b)
        o = self._opts(tol, 20, maxit, False)
        info = _lib.SolveInfo()
        with torch.cuda.device(self.device):
            self._bind()
            code = fn(self._ctx, int(mode), ctypes.c_void_p(_dptr(b)), ctypes.c_void_p(_dptr(x)), ctypes.byref(o),
                      ctypes.byref(info))
        check(code, allow=() if raise_on_notconv else (_lib.HELM3D_ERR_NOTCONV,))
        return x, SolveInfo(int(info.iters), int(info.cycles), float(info.relres), float(info.quasi_relres),
                            bool(info.converged), inner_iters=int(info.inner_iters))

    def solve_crmn(self, b: torch.Tensor, mode: int = 1, tol: float = 1e-6, maxit: int = 10000,
                   raise_on_notconv: bool = False):
        """CRMN.m on the Kaczmarz system (linearsolve.m:52-63, omega 1.5)."""
        return self.solve_cgmn(b, mode, tol, maxit, raise_on_notconv, method="crmn")

    def set_graphs(self, enable: bool):
        """Replay fixed kernel sequences from captured HIP graphs (default on)."""
        with torch.cuda.device(self.device):
            check(lib().helm3d_set_option(self._ctx, _lib.HELM3D_OPT_GRAPHS, int(enable)))

    # -- launch shape ---------------------------------------------------------
    def set_launch(self, kchunk: int = 0, tile_rows: int = 4, prefetch: int = 1, rows_per_thread: int = 1):
        """Stencil launch shape (results are identical for every shape)."""
        with torch.cuda.device(self.device):
            check(lib().helm3d_set_launch(self._ctx, int(kchunk), int(tile_rows), int(prefetch),
                                          int(rows_per_thread)))

    def autotune(self, mode: int = 1, precision: str = "c128", reps: int = 10) -> dict:
        """Pick the fastest stencil launch shape for this grid (helm3d_autotune)."""
        prec = {"c128": _lib.HELM3D_C128, "c64": _lib.HELM3D_C64}[precision]
        v = [ctypes.c_int(0) for _ in range(4)]
        ms = ctypes.c_double(0)
        with torch.cuda.device(self.device):
            self._bind()
            check(lib().helm3d_autotune(self._ctx, int(mode), prec, int(reps), *[ctypes.byref(a) for a in v],
                                        ctypes.byref(ms)))
        return dict(kchunk=v[0].value, tile_rows=v[1].value, prefetch=v[2].value, rows_per_thread=v[3].value,
                    ms=ms.value)

    def get_launch(self) -> dict:
        v = [ctypes.c_int(0) for _ in range(5)]
        with torch.cuda.device(self.device):
            check(lib().helm3d_get_launch(self._ctx, *[ctypes.byref(a) for a in v]))
        return dict(kchunk=v[0].value, tile_rows=v[1].value, prefetch=v[2].value, rows_per_thread=v[3].value,
                    blocks=v[4].value)

    # -- in-library kernel timing -------------------------------------------
    def profile(self, enable: bool = True):
        """Enable/disable HIP-event timing of every kernel class (resets totals)."""
        with torch.cuda.device(self.device):
            self._bind()
            check(lib().helm3d_profile(self._ctx, int(enable)))

    def profile_read(self) -> dict:
        """{kind: (total_ms, launches)} since the last profile() call."""
        out = {}
        with torch.cuda.device(self.device):
            self._bind()
            for name, k in _lib.PROF_KINDS.items():
                ms = ctypes.c_double(0)
                n = ctypes.c_longlong(0)
                check(lib().helm3d_profile_read(self._ctx, k, ctypes.byref(ms), ctypes.byref(n)))
                out[name] = (ms.value, n.value)
        return out

    # -- gradient -----------------------------------------------------------
    def grad_accum(self, U: torch.Tensor, V: torch.Tensor, g: torch.Tensor, scale: float = 1.0,
                   dwn: Optional[torch.Tensor] = None, grid: str = "phys"):
        """g += scale * Re(conj(U) .* (dH^H V)) (PDEfunc.m:394 before comp2phys).
        grid "phys": g on the physical grid, PML truncated (cut_pml);
        grid "pml": g on the extended grid (fold with extension_adjoint)."""
        U = self._vec(U, torch.complex128)
        V = self._vec(V, torch.complex128)
        gcode = {"phys": _lib.HELM3D_GRID_PHYS, "pml": _lib.HELM3D_GRID_PML}[grid]
        size = int(np.prod(self.n_phys)) if gcode == _lib.HELM3D_GRID_PHYS else self.N
        if (g.dtype != torch.float64 or not g.is_contiguous() or g.numel() != size or g.device != self.device):
            raise ValueError(f"g must be a contiguous float64 tensor of {size} elements on {self.device}")
        d = None
        if dwn is not None:
            d = dwn.reshape(-1).to(device=self.device, dtype=torch.float64).contiguous()
            if d.numel() != self.N:
                raise ValueError("dwn must have prod(n) elements")
        with torch.cuda.device(self.device):
            self._bind()
            check(lib().helm3d_fwi_grad_accum(self._ctx, ctypes.c_void_p(_dptr(d)) if d is not None else None,
                                              ctypes.c_void_p(_dptr(U)), ctypes.c_void_p(_dptr(V)),
                                              float(scale), gcode, ctypes.c_void_p(_dptr(g))))
        return g

    # -- model set-up and sparse source / receiver operators ------------------
    def param2wavenum(self, v: torch.Tensor, n_phys, npml, freq: float, unit: str = "m/s",
                      derivs: int = 1):
        """[wn, dwn, ddwn] = param2wavenum(Pext v, freq, unit) on the extended
        grid (discrete_helmholtz.m:240-244, param2wavenum.m:20-49).  Returns
        (wn, dwn or None, ddwn or None) as float64 device tensors."""
        n_phys = _as_int3(n_phys)
        p6 = npml_flat(npml)
        if unit not in _lib.HELM3D_UNITS:
            raise ValueError("Unknown unit - " + str(unit))
        v = v.reshape(-1).to(device=self.device, dtype=torch.float64).contiguous()
        if v.numel() != int(np.prod(n_phys)):
            raise ValueError("v must have prod(n_phys) elements")
        nt = tuple(n_phys[a] + p6[2 * a] + p6[2 * a + 1] for a in range(3))
        N = int(np.prod(nt))
        wn = torch.empty(N, dtype=torch.float64, device=self.device)
        dwn = torch.empty_like(wn) if derivs >= 1 else None
        ddwn = torch.empty_like(wn) if derivs >= 2 else None
        P = lambda t: ctypes.c_void_p(_dptr(t)) if t is not None else None  # noqa: E731
        with torch.cuda.device(self.device):
            self._bind()
            check(lib().helm3d_param2wavenum(self._ctx, (ctypes.c_int * 3)(*n_phys), (ctypes.c_int * 6)(*p6),
                                             P(v), _lib.HELM3D_UNITS[unit], float(freq), P(wn), P(dwn), P(ddwn)))
        return wn, dwn, ddwn

    def extend(self, t: torch.Tensor, n_phys, npml) -> torch.Tensor:
        """Pext t: edge replication of a physical-grid field into the PML
        (opExtension.m:94-96); param2wavenum with c = (2 pi f)^2 = 1, s2/m2."""
        return self.param2wavenum(t, n_phys, npml, 1.0 / (2.0 * np.pi), "s2/m2", derivs=0)[0]

    def extension_adjoint(self, gpml: torch.Tensor, n_phys, npml, g: Optional[torch.Tensor] = None,
                          accumulate: bool = False) -> torch.Tensor:
        """g (+)= Pext' gpml: adjoint of the edge replication (opExtension.m:97-101)."""
        n_phys = _as_int3(n_phys)
        p6 = npml_flat(npml)
        nt = tuple(n_phys[a] + p6[2 * a] + p6[2 * a + 1] for a in range(3))
        if gpml.dtype != torch.float64 or not gpml.is_contiguous() or gpml.numel() != int(np.prod(nt)):
            raise ValueError("gpml must be a contiguous float64 tensor on the extended grid")
        if g is None:
            g = torch.empty(int(np.prod(n_phys)), dtype=torch.float64, device=self.device)
            accumulate = False
        elif g.dtype != torch.float64 or not g.is_contiguous() or g.numel() != int(np.prod(n_phys)):
            raise ValueError("g must be a contiguous float64 tensor on the physical grid")
        with torch.cuda.device(self.device):
            self._bind()
            check(lib().helm3d_extension_adjoint(self._ctx, (ctypes.c_int * 3)(*n_phys), (ctypes.c_int * 6)(*p6),
                                                 ctypes.c_void_p(_dptr(gpml)), ctypes.c_void_p(_dptr(g)),
                                                 int(accumulate)))
        return g

    def csr_apply(self, op: "SparseRows", x: torch.Tensor, out: torch.Tensor, alpha: float = 1.0,
                  accumulate: bool = False) -> torch.Tensor:
        """out[rows] (+)= alpha * (op x)[rows] for a SparseRows operator; rows of
        op without entries are left untouched."""
        if x.dtype != torch.complex128 or out.dtype != torch.complex128:
            raise ValueError("csr_apply works on complex128 vectors")
        if x.numel() != op.shape[1] or out.numel() != op.shape[0]:
            raise ValueError(f"shape mismatch: op {op.shape}, x {x.numel()}, out {out.numel()}")
        if not (x.is_contiguous() and out.is_contiguous()):
            raise ValueError("vectors must be contiguous")
        with torch.cuda.device(self.device):
            self._bind()
            check(lib().helm3d_csr_apply(self._ctx, op.nrows, ctypes.c_void_p(_dptr(op.rowptr)),
                                         ctypes.c_void_p(_dptr(op.col)), ctypes.c_void_p(_dptr(op.w)),
                                         ctypes.c_void_p(_dptr(op.oidx)), float(alpha),
                                         ctypes.c_void_p(_dptr(x)), ctypes.c_void_p(_dptr(out)),
                                         int(accumulate)))
        return out


class SparseRows:
    """A real sparse matrix (shape m x n) kept on the device as its non-empty
    rows: rowptr/col/w in row-compressed form and oidx, the output row of each
    stored row.  Built from a scipy.sparse matrix."""

    def __init__(self, A, device):
        import scipy.sparse as sp
        A = sp.csr_matrix(A, dtype=np.float64)
        A.sum_duplicates()
        A.eliminate_zeros()
        self.shape = A.shape
        counts = np.diff(A.indptr)
        rows = np.nonzero(counts)[0]
        sub = A[rows]
        self.nrows = int(rows.size)
        dev = torch.device(device)
        self.rowptr = torch.from_numpy(sub.indptr.astype(np.int64)).to(dev)
        self.col = torch.from_numpy(sub.indices.astype(np.int64)).to(dev)
        self.w = torch.from_numpy(sub.data.astype(np.float64)).to(dev)
        self.oidx = torch.from_numpy(rows.astype(np.int64)).to(dev)
        self.host = A


# ---------------------------------------------------------------------------
# Drop-in matvec at the MEX boundary (host numpy arrays)
# ---------------------------------------------------------------------------

_VARIANT_NAMES = {
    "forw": _lib.HELM3D_FORW, "adj": _lib.HELM3D_ADJ,
    "forw_deriv": _lib.HELM3D_FORW | _lib.HELM3D_DERIV, "adj_deriv": _lib.HELM3D_ADJ | _lib.HELM3D_DERIV,
    "forw_wnc": _lib.HELM3D_FORW | _lib.HELM3D_WNC, "adj_wnc": _lib.HELM3D_ADJ | _lib.HELM3D_WNC,
}


def helm3d_mvp_host(variant, wn, h, n, npml, x, nthreads: int = 1) -> np.ndarray:
    """y = Helm3dmvp_<variant>_mex(wn, h, n, npml, x, nthreads) on host arrays
    (split-complex boundary of Helm3dmvp_mex.c:58-107), computed on the GPU."""
    v = _VARIANT_NAMES[variant] if isinstance(variant, str) else int(variant)
    wn = np.asarray(wn).reshape(-1)
    n_arr = np.ascontiguousarray(np.asarray(n, dtype=np.float64).reshape(-1))
    x = np.asarray(x)
    x_rows = x.shape[0] if x.ndim >= 1 else 1
    x = x.reshape(x_rows, -1)[:, 0] if x.ndim > 1 else x.reshape(-1)
    N = x.size
    wnr = np.ascontiguousarray(wn.real, dtype=np.float64)
    wni = np.ascontiguousarray(wn.imag, dtype=np.float64) if np.iscomplexobj(wn) else None
    xr = np.ascontiguousarray(x.real, dtype=np.float64)
    xi = np.ascontiguousarray(x.imag, dtype=np.float64) if np.iscomplexobj(x) else None
    hh = np.ascontiguousarray(np.asarray(h, dtype=np.float64).reshape(3))
    pp = np.ascontiguousarray(np.asarray(npml_flat(npml), dtype=np.float64))
    yr = np.zeros(N)
    yi = np.zeros(N)
    P = lambda a: a.ctypes.data_as(_lib._dp) if a is not None else None  # noqa: E731
    code = lib().helm3d_mvp(v, P(wnr), P(wni), P(hh), P(n_arr), n_arr.size, P(pp), P(yr), P(yi), P(xr), P(xi),
                            x_rows, int(nthreads))
    check(code)
    return yr + 1j * yi


_SWEEP_VARIANTS = {"forw": _lib.HELM3D_FORW, "adj": _lib.HELM3D_ADJ,
                   "forw_wnc": _lib.HELM3D_FORW | _lib.HELM3D_WNC, "adj_wnc": _lib.HELM3D_ADJ | _lib.HELM3D_WNC}


def helm3d_kaczswp_host(variant, wn, h, n, npml, omega, x, b) -> np.ndarray:
    """y = Helm3d_kacz_sweep{,_wnc,_adj}_mex(wn, h, n, npml, w, x, b) on host
    arrays (27-colour double sweep on the GPU).  x and b are N-vectors or
    N x P matrices; column s of y sweeps column s of x against column s of b
    (Helm3dkaczswp_mex.c:159,182-186).  Returns an N-vector for vector x,
    else N x P."""
    v = _SWEEP_VARIANTS[variant] if isinstance(variant, str) else int(variant)
    wn = np.asarray(wn).reshape(-1)
    n_arr = np.ascontiguousarray(np.asarray(n, dtype=np.float64).reshape(-1))
    x = np.asarray(x)
    b = np.asarray(b)
    vec_out = x.ndim <= 1
    x2 = x.reshape(-1, 1) if x.ndim <= 1 else x.reshape(x.shape[0], -1)
    b2 = b.reshape(-1, 1) if b.ndim <= 1 else b.reshape(b.shape[0], -1)
    x_rows, x_cols = x2.shape
    b_rows, b_cols = b2.shape
    # column-major N x P blocks, as mxGetPr delivers them
    f = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64).T).reshape(-1)  # noqa: E731
    wnr = np.ascontiguousarray(wn.real, dtype=np.float64)
    wni = np.ascontiguousarray(wn.imag, dtype=np.float64) if np.iscomplexobj(wn) else None
    xr, xi = f(x2.real), (f(x2.imag) if np.iscomplexobj(x2) else None)
    br, bi = f(b2.real), (f(b2.imag) if np.iscomplexobj(b2) else None)
    hh = np.ascontiguousarray(np.asarray(h, dtype=np.float64).reshape(3))
    pp = np.ascontiguousarray(np.asarray(npml_flat(npml), dtype=np.float64))
    yr = np.zeros(x_rows * x_cols)
    yi = np.zeros(x_rows * x_cols)
    P = lambda a: a.ctypes.data_as(_lib._dp) if a is not None else None  # noqa: E731
    code = lib().helm3d_kaczswp(v, P(wnr), P(wni), P(hh), P(n_arr), n_arr.size, P(pp), float(omega), P(xr), P(xi),
                                x_rows, x_cols, P(br), P(bi), b_rows, b_cols, P(yr), P(yi))
    check(code, ids=_lib.MEX_SWEEP_ERR_IDS)
    y = (yr + 1j * yi).reshape(x_cols, x_rows).T
    return y[:, 0] if vec_out else y
