"""3D frequency-domain FWI on the device: the PDEfunc modes on the Helmholtz solver.

Mirrors tools/algorithms/CommonFreqModeling/PDEfunc.m for the 3D
``operto27`` scheme with the matrix-free operator:

* ``discrete_helmholtz`` (discrete_helmholtz.m:105-244): PML widths, the
  extended grid, ``wn``/``dwn`` from the model (one device kernel: edge
  replication + param2wavenum);
* sources and receivers (DataGrid.m:22-44, RegularGrid.m:45-69,
  opInterp.m:34-57, opSincInterp.m:26-51, kaiser_window.m:17-19): separable
  sinc/Kaiser weights built on the host once, applied on the device as
  row-compressed gathers;
* the source term ``Qk = Ps (w(f) Q) prod(d)/prod(dt)`` with the Ricker
  spectrum of fwi_wavelet.m:24-28 (PDEfunc.m:207,324-336);
* ``OBJ`` (PDEfunc.m:383-394): f = sum 1/2 ||Pr U - D||^2 (LS_misfit.m:21-26),
  g = comp2phys sum Re(conj(U) .* dH^H V) with V = H^-H (-Pr^T (Pr U - D));
* ``FORW_MODEL``, ``JACOB_FORW``, ``JACOB_ADJ``, ``HESS_GN``
  (PDEfunc.m:421-444) and the full Hessian-vector product ``HESS``
  (PDEfunc.m:444-450, discrete_helmholtz.m:286-293);
* ``zero_boundary`` (PDEfunc.m:478-482).

Every solve, stencil, gradient, injection and sampling runs in the HIP
library; torch holds the device buffers and does the per-receiver residual
arithmetic (nrec values per source).  ``FWIObjective`` wraps OBJ as a
``torch.autograd.Function`` of the model.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from .core import Helm3d, SparseRows

OBJ = "obj"
FORW_MODEL = "forw_model"
JACOB_FORW = "jacob_forw"
JACOB_ADJ = "jacob_adj"
HESS_GN = "hess_gn"
HESS = "hess"
MODES = (OBJ, FORW_MODEL, JACOB_FORW, JACOB_ADJ, HESS_GN, HESS)

# Kaiser window parameter b by half-width r (opSincInterp.m:41-43); the
# source and receiver operators use r = 4 (opInterp.m:34).
KAISER_B = (1.24, 2.94, 4.53, 6.31, 7.91, 9.42, 10.95, 12.53, 14.09, 14.18)
SINC_WINDOW = 4


@dataclass
class FWIModel:
    """The ``model`` struct of PDEfunc.m:25-33 (3D): grid o/d/n (x, y, z),
    frequencies, wavelet (f0, t0), unit and the source / receiver coordinate
    vectors (each set is the tensor grid of its three vectors, x fastest)."""
    o: Sequence[float]
    d: Sequence[float]
    n: Sequence[int]
    freq: Sequence[float]
    xsrc: Sequence[float]
    ysrc: Sequence[float]
    zsrc: Sequence[float]
    xrec: Sequence[float]
    yrec: Sequence[float]
    zrec: Sequence[float]
    f0: float = 10.0
    t0: float = 0.0
    unit: str = "s2/m2"

    @property
    def nsrc(self) -> int:
        return len(self.xsrc) * len(self.ysrc) * len(self.zsrc)

    @property
    def nrec(self) -> int:
        return len(self.xrec) * len(self.yrec) * len(self.zrec)


@dataclass
class FWIOpts:
    """PDEopts.m:108-127 and LinSolveOpts.m:55-62 fields used on this path.

    solver: "fgmres" (FGMRES.m in complex128), "mixed" (FP64 flexible GMRES
    with a complex64 BiCGStab preconditioner) or "mg" (FGMRES preconditioned
    by the ML_GMRES V-cycle, mg_levels levels, mg_k = ML_GMRES.m's
    (ks_outer, ks_inner, kc_outer, kc_inner), the V-cycle in complex128 or,
    with mg_precision "mixed", in complex64; mg_workers > 1 solves that many
    right-hand sides at once on separate streams); tol is the relative
    residual of every solve.  autotune picks the stencil launch shape per frequency
    (one-at-a-time solves).
    batch > 1 solves that many sources of a frequency together
    (helm3d_batch_solve: FGMRES_MT.m semantics for "fgmres", a fixed
    inner_iters complex64 BiCGStab iterations per outer step for "mixed")."""
    pml: Optional[object] = 10
    pml_max: int = 10
    cut_pml: bool = True
    free_surface: bool = False
    zero_boundary: bool = False
    src_interp: str = "sinc"
    rec_interp: str = "sinc"
    solver: str = "fgmres"
    tol: float = 1e-6
    maxinnerit: int = 20
    maxit: int = 10000
    inner_tol: float = 1e-2
    inner_maxit: int = 200
    batch: int = 1
    inner_iters: int = 20
    autotune: bool = False
    mg_levels: int = 3
    mg_k: tuple = (3, 5, 3, 5)
    mg_precision: str = "c128"
    mg_workers: int = 1
    stats: list = field(default_factory=list, repr=False)


def fwi_wavelet(f, t0: float, f0: float):
    """fwi_wavelet.m:24-28: w = f^2 exp(-(f/f0)^2) exp(-2 pi i f t0)."""
    f = np.asarray(f, dtype=np.float64)
    w = np.exp(-1j * 2 * np.pi * f * t0)
    if f0:
        w = f ** 2 * np.exp(-(f / f0) ** 2) * w
    return w


class FWIFreqData:
    """Frequency-domain data with its acquisition metadata (FWIFreqData.m:27-48,
    104-113, numeric-data form): nrec x (nsrc nfreq) with frequencies outer,
    as PDEfunc indexes it (column k*nsrc + j for frequency k, source j)."""

    def __init__(self, D, model: "FWIModel"):
        self.freqgrid = list(model.freq)
        self.model = model
        D = D if isinstance(D, torch.Tensor) else torch.as_tensor(np.asarray(D))
        nrec = model.nrec
        if D.numel() % nrec:
            raise ValueError("data size is not a multiple of the receiver count")
        if D.dim() == 2 and D.shape[0] == nrec:
            self.localData = D
        else:
            # reshape(D, nrec, []) of a column-major (MATLAB) array
            self.localData = D.reshape(-1).reshape(D.numel() // nrec, nrec).T

    def getData(self, index=None):
        return self.localData if index is None else self.localData[:, index]

    def setData(self, index, data):
        self.localData[:, index] = torch.as_tensor(data, dtype=self.localData.dtype,
                                                   device=self.localData.device).reshape(self.localData.shape[0], -1)

    def getSourceGrid(self):
        m = self.model
        return dict(src=(m.xsrc, m.ysrc, m.zsrc), rec=(m.xrec, m.yrec, m.zrec))


def grid_coords(o: float, d: float, n: int) -> np.ndarray:
    return o + d * np.arange(n)


def lagrange_cubic(x1, x2) -> np.ndarray:
    """Weights (len(x2) x len(x1)) of opLInterp1D(x1, x2) (opLInterp1D.m:114-153):
    4-point Lagrange interpolation from the grid x1 to the points x2, with the
    bracketing window walked from point to point as the reference does; points
    outside [x1(1), x1(end)] get a zero row.  Needs at least 4 points in x1."""
    x1 = np.asarray(x1, dtype=np.float64).reshape(-1)
    x2 = np.asarray(x2, dtype=np.float64).reshape(-1)
    n1 = x1.size
    W = np.zeros((x2.size, n1))
    ik = np.nonzero((x2 >= x1[0]) & (x2 <= x1[-1]))[0]
    if ik.size == 0:
        return W
    if n1 < 4:
        raise ValueError("cubic interpolation needs at least 4 grid points")
    a, b, c, d = 0, 1, 2, 3  # 0-based window
    for k in ik:
        x = x2[k]
        if x < x1[b]:
            while x < x1[b] and b - 1 > 0:
                b -= 1
            a, c = b - 1, b + 1
            d = c + 1
        elif x > x1[c]:
            while x > x1[c] and c + 1 < n1 - 1:
                c += 1
            a, b, d = c - 2, c - 1, c + 1
        idx = (a, b, c, d)
        for q in range(4):
            num = 1.0
            den = 1.0
            for r in range(4):
                if r != q:
                    num *= x - x1[idx[r]]
                    den *= x1[idx[q]] - x1[idx[r]]
            W[k, idx[q]] += num / den
    return W


def interp_1d(xin, xout, kind: str = "sinc") -> np.ndarray:
    """Weights (len(xout) x len(xin)) of the 1D interpolation of opInterp.m:41-47:
    "sinc": opSincInterp(xin, xout, 4), the Kaiser-windowed sinc of
    (xout - xin)/dx with dx the xout spacing (opSincInterp.m:26-51,
    kaiser_window.m:17-19); "cubic": opLInterp1D(xin, xout)."""
    if kind == "cubic":
        return lagrange_cubic(xin, xout)
    if kind != "sinc":
        raise ValueError("Unrecognized type")
    xin = np.asarray(xin, dtype=np.float64).reshape(-1)
    xout = np.asarray(xout, dtype=np.float64).reshape(-1)
    if xout.size > 1:
        dx = xout[1] - xout[0]
        xin = xin / dx
        xout = xout / dx
    X = xout[:, None] - xin[None, :]
    r = SINC_WINDOW
    b = KAISER_B[r - 1]
    win = np.zeros_like(X)
    inside = np.abs(X) <= r
    win[inside] = np.i0(b * np.sqrt(1.0 - (X[inside] / r) ** 2)) / np.i0(b)
    W = win * np.sinc(X)
    # Points that sit on the grid up to rounding of (xout - xin)/dx give
    # weights of order 1e-16 at the integer offsets; dropping them keeps the
    # operator a unit impulse there (the relative change is below 1e-13).
    W[np.abs(W) < 1e-13] = 0.0
    return W


def interp_3d(pts_x, pts_y, pts_z, comp_x, comp_y, comp_z, kind: str = "sinc"):
    """scipy.sparse (ncomp x npts) of opKron(Iz, Iy, Ix): interpolation from
    the tensor grid (pts_x, pts_y, pts_z) onto the computational grid
    (opInterp.m:51-57, RegularGrid.m:55-56 in adjoint mode)."""
    import scipy.sparse as sp
    Wx = sp.csr_matrix(interp_1d(pts_x, comp_x, kind))
    Wy = sp.csr_matrix(interp_1d(pts_y, comp_y, kind))
    Wz = sp.csr_matrix(interp_1d(pts_z, comp_z, kind))
    return sp.kron(Wz, sp.kron(Wy, Wx, format="csr"), format="csr")


def pml_widths(v_phys: np.ndarray, model: FWIModel, freq: float, opts: FWIOpts) -> np.ndarray:
    """npml as the 2x3 [lo;hi] matrix of discrete_helmholtz.m:169-201."""
    dt = np.asarray(model.d, dtype=np.float64)
    if opts.pml is not None and np.size(opts.pml) > 0:
        p = np.asarray(opts.pml).reshape(-1)
        npml = np.full(3, int(p[0])) if p.size == 1 else p.astype(int)
        if npml.size != 3:
            raise ValueError("npml must have length 2 or 3")
    else:
        vmax = {"m/s": lambda v: v.max(), "s2/m2": lambda v: (v ** -0.5).max(),
                "s2/km2": lambda v: (1e3 * v ** -0.5).max()}[model.unit](v_phys)
        npml = np.ceil(vmax / (np.real(freq) * dt)).astype(int)
    npml = np.minimum(npml, opts.pml_max)
    npml = np.vstack([npml, npml])
    if opts.free_surface:
        npml[0, 2] = 0
    return npml


class FWIProblem:
    """Device-resident state of the PDEfunc computations for one model grid,
    source / receiver geometry and solver choice."""

    def __init__(self, model: FWIModel, opts: Optional[FWIOpts] = None, Q=None, device=None):
        self.model = model
        self.opts = opts or FWIOpts()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.n = tuple(int(a) for a in model.n)
        self.d = tuple(float(a) for a in model.d)
        self.Q = np.eye(model.nsrc) if Q is None else np.asarray(Q)
        if self.Q.shape[0] != model.nsrc:
            raise ValueError("size(Q,1) must match the source grid")
        self.H = Helm3d(device=self.device)
        self._batch = None
        self._mg = None   # multigrid hierarchy of the current frequency (solver "mg")
        self._geom = {}   # npml key -> (Ps, Pr, Pr^T)

    # -- grids and operators --------------------------------------------------
    def _geometry(self, npml: np.ndarray):
        key = tuple(npml.reshape(-1))
        if key not in self._geom:
            m = self.model
            nt = [self.n[a] + int(npml[0, a] + npml[1, a]) for a in range(3)]
            ot = [m.o[a] - npml[0, a] * self.d[a] for a in range(3)]
            comp = [grid_coords(ot[a], self.d[a], nt[a]) for a in range(3)]
            Ps = interp_3d(m.xsrc, m.ysrc, m.zsrc, *comp, kind=self.opts.src_interp)
            Wr = interp_3d(m.xrec, m.yrec, m.zrec, *comp, kind=self.opts.rec_interp)
            # Pr = (interpolation receivers -> grid)' (DataGrid.m:39)
            self._geom[key] = (SparseRows(Ps, self.device), SparseRows(Wr.T.tocsr(), self.device),
                               SparseRows(Wr, self.device))
        return self._geom[key]

    def _setup(self, v: torch.Tensor, freq: float):
        """discrete_helmholtz for one frequency: npml, wn, dwn; the operator is
        loaded into the context."""
        m = self.model
        v_host = None
        if self.opts.pml is None:
            v_host = v.detach().cpu().numpy()
        npml = pml_widths(v_host, m, freq, self.opts)
        p6 = tuple(int(a) for a in npml.T.reshape(-1))
        wn, dwn, _ = self.H.param2wavenum(v, self.n, p6, freq, m.unit, derivs=1)
        nt = tuple(self.n[a] + p6[2 * a] + p6[2 * a + 1] for a in range(3))
        self.H.set_model(nt, self.d, p6, wn)
        if self.opts.autotune and self.opts.batch <= 1:
            # launch shape of the stencil the one-at-a-time solves spend their time in
            self.H.autotune(mode=1, precision="c64" if self.opts.solver == "mixed" else "c128", reps=5)
        if self.opts.solver == "mg":
            key = (nt, p6, self.opts.mg_levels, tuple(self.opts.mg_k), self.opts.mg_workers)
            if self._mg is None or self._mg[0] != key:
                from .mg import MGPool
                self._mg = None
                self._mg = (key, MGPool(v, self.n, self.d, int(npml[0, 0]), freq, m.unit, self.opts.mg_levels,
                                        self.opts.mg_k, self.device, free_surface=self.opts.free_surface,
                                        workers=self.opts.mg_workers))
            else:
                self._mg[1].set_model(v, freq)
        Ps, Pr, PrT = self._geometry(npml)
        return dict(npml=npml, p6=p6, nt=nt, N=int(np.prod(nt)), wn=wn, dwn=dwn, Ps=Ps, Pr=Pr, PrT=PrT)

    def _dwn_args(self, S, freq: float):
        """(dwn tensor or None, scale) for the DERIV stencil: a constant dwn
        (s2/m2, s2/km2) is passed as the scale of a unit dwn."""
        if self.model.unit == "m/s":
            return S["dwn"], 1.0
        c = (2 * np.pi * freq) ** 2
        return None, c * (1e-6 if self.model.unit == "s2/km2" else 1.0)

    def _dwn_times(self, S, freq: float, dm: torch.Tensor) -> torch.Tensor:
        """dwn .* (Pext dm) on the extended grid: T(U) dm = dH (dm .* U) is the
        DERIV stencil with this coefficient applied to U."""
        dm_pml = self.H.extend(dm, self.n, S["p6"])
        dwn, scale = self._dwn_args(S, freq)
        return dm_pml * scale if dwn is None else dm_pml * dwn

    def pair_bytes(self, v: torch.Tensor) -> int:
        """Device bytes of one complex128 wavefield per (frequency, source)."""
        v_host = v.detach().cpu().numpy() if self.opts.pml is None else None
        nt_max = 0
        for f in self.model.freq:
            npml = pml_widths(v_host, self.model, f, self.opts)
            nt_max = max(nt_max, int(np.prod([self.n[a] + int(npml[:, a].sum()) for a in range(3)])))
        return 16 * nt_max

    def _solve(self, b: torch.Tensor, mode: int) -> torch.Tensor:
        o = self.opts
        if o.solver == "fgmres":
            x, info = self.H.solve(b, mode=mode, tol=o.tol, maxinnerit=o.maxinnerit, maxit=o.maxit)
        elif o.solver == "mixed":
            x, info = self.H.solve_mixed(b, mode=mode, tol=o.tol, maxinnerit=o.maxinnerit, maxit=o.maxit,
                                         inner_tol=o.inner_tol, inner_maxit=o.inner_maxit)
        elif o.solver == "mg":
            xs, infos = self._mg[1].solve_many([b], mode=mode, tol=o.tol, maxinnerit=o.maxinnerit, maxit=o.maxit,
                                               precision=o.mg_precision)
            x, info = xs[0], infos[0]
        else:
            raise ValueError("solver must be 'fgmres', 'mixed' or 'mg'")
        o.stats.append(info)
        return x

    def _solve_many(self, bs, mode: int):
        """Solves for a list of right-hand sides, opts.batch at a time."""
        o = self.opts
        if o.solver == "mg":
            xs, infos = self._mg[1].solve_many(bs, mode=mode, tol=o.tol, maxinnerit=o.maxinnerit, maxit=o.maxit,
                                               precision=o.mg_precision)
            o.stats.extend(infos)
            return xs
        if o.batch <= 1:
            return [self._solve(b, mode) for b in bs]
        from .batch import BatchSolver
        if self._batch is None:
            self._batch = BatchSolver(self.device)
        method = {"fgmres": "plain", "mixed": "mixed"}[o.solver]
        out = []
        for i in range(0, len(bs), o.batch):
            chunk = bs[i:i + o.batch]
            xs, infos = self._batch.solve([self.H] * len(chunk), chunk, mode=mode, method=method, tol=o.tol,
                                          maxinnerit=o.maxinnerit, maxit=o.maxit, inner_maxit=o.inner_iters)
            o.stats.extend(infos)
            out.extend(xs)
        return out

    def _source(self, S, k: int, isrc: int) -> torch.Tensor:
        """Qk_i = Ps (w(f) Q(:, isrc)) prod(d)/prod(dt) (PDEfunc.m:324-336);
        dt = d on this path, so the ratio is 1."""
        m = self.model
        w = fwi_wavelet([m.freq[k]], m.t0, m.f0)[0]
        q = torch.from_numpy(np.ascontiguousarray(w * self.Q[:, isrc].astype(np.complex128))).to(self.device)
        b = torch.zeros(S["N"], dtype=torch.complex128, device=self.device)
        self.H.csr_apply(S["Ps"], q, b)
        return b

    def _sample(self, S, U: torch.Tensor) -> torch.Tensor:
        out = torch.zeros(self.model.nrec, dtype=torch.complex128, device=self.device)
        self.H.csr_apply(S["Pr"], U, out)
        return out

    def _inject_rec(self, S, r: torch.Tensor, alpha: float) -> torch.Tensor:
        b = torch.zeros(S["N"], dtype=torch.complex128, device=self.device)
        self.H.csr_apply(S["PrT"], r.to(torch.complex128).contiguous(), b, alpha=alpha)
        return b

    def _grad_buffers(self, S):
        if self.opts.cut_pml:
            return None
        return torch.zeros(S["N"], dtype=torch.float64, device=self.device)

    def _grad_term(self, S, freq, U, V, g_phys, gpml, weight: float = 1.0):
        dwn, scale = self._dwn_args(S, freq)
        if gpml is None:
            self.H.grad_accum(U, V, g_phys, scale=scale * weight, dwn=dwn, grid="phys")
        else:
            self.H.grad_accum(U, V, gpml, scale=scale * weight, dwn=dwn, grid="pml")

    def _grad_finish(self, S, g_phys, gpml):
        if gpml is not None:
            self.H.extension_adjoint(gpml, self.n, S["p6"], g=g_phys, accumulate=True)

    def _pairs(self):
        """(frequency index, source index, data column) in PDEfunc.m order:
        frequencies outer, sources inner (PDEfunc.m:252-262,322-330)."""
        nsrc = self.Q.shape[1]
        for k in range(len(self.model.freq)):
            for j in range(nsrc):
                yield k, j, k * nsrc + j

    def _zero_boundary(self, g: torch.Tensor) -> torch.Tensor:
        if self.opts.zero_boundary:
            G = g.view(self.n[2], self.n[1], self.n[0])
            G[[0, -1], :, :] = 0
            G[:, [0, -1], :] = 0
            G[:, :, [0, -1]] = 0
        return g

    def _model_tensor(self, v) -> torch.Tensor:
        if isinstance(v, torch.Tensor):
            t = v.detach().to(device=self.device, dtype=torch.float64)
        else:
            t = torch.from_numpy(np.ascontiguousarray(np.asarray(v, dtype=np.float64))).to(self.device)
        t = t.reshape(-1).contiguous()
        if t.numel() != int(np.prod(self.n)):
            raise ValueError("v must have prod(model.n) elements")
        return t

    def _data(self, Dobs) -> torch.Tensor:
        if isinstance(Dobs, FWIFreqData):
            Dobs = Dobs.getData()
        D = Dobs if isinstance(Dobs, torch.Tensor) else torch.from_numpy(np.asarray(Dobs))
        D = D.to(device=self.device, dtype=torch.complex128)
        npairs = len(self.model.freq) * self.Q.shape[1]
        D = D.reshape(self.model.nrec, npairs)
        return D

    # -- PDEfunc modes ----------------------------------------------------------
    def forward_wavefields(self, v, keep: bool = True):
        """Forward solves for every (frequency, source): (predicted data
        nrec x npairs, list of U or None)."""
        v = self._model_tensor(v)
        nsrc = self.Q.shape[1]
        pred = torch.zeros(self.model.nrec, len(self.model.freq) * nsrc, dtype=torch.complex128,
                           device=self.device)
        Us = []
        for k, freq in enumerate(self.model.freq):
            S = self._setup(v, freq)
            Uk = self._solve_many([self._source(S, k, j) for j in range(nsrc)], 1)
            for j, U in enumerate(Uk):
                pred[:, k * nsrc + j] = self._sample(S, U)
                Us.append(U if keep else None)
        return pred, Us

    def objective(self, v, Dobs, need_grad: bool = True):
        """PDEfunc(OBJ): (f, g) with g on the physical grid (float64 device
        tensor), or (f, None)."""
        v = self._model_tensor(v)
        D = self._data(Dobs)
        nsrc = self.Q.shape[1]
        f = 0.0
        g = torch.zeros(int(np.prod(self.n)), dtype=torch.float64, device=self.device) if need_grad else None
        for k, freq in enumerate(self.model.freq):
            S = self._setup(v, freq)
            gpml = self._grad_buffers(S) if need_grad else None
            Us = self._solve_many([self._source(S, k, j) for j in range(nsrc)], 1)
            rs = []
            for j, U in enumerate(Us):
                r = self._sample(S, U) - D[:, k * nsrc + j]
                f += 0.5 * float(torch.vdot(r, r).real)
                rs.append(r)
            if need_grad:
                Vs = self._solve_many([self._inject_rec(S, r, -1.0) for r in rs], 2)
                for U, V in zip(Us, Vs):
                    self._grad_term(S, freq, U, V, g, gpml)
                self._grad_finish(S, g, gpml)
        if need_grad:
            self._zero_boundary(g)
        return f, g

    def forw_model(self, v) -> torch.Tensor:
        """PDEfunc(FORW_MODEL): Pr U for every (frequency, source), nrec x npairs."""
        return self.forward_wavefields(v, keep=False)[0]

    def jacob_forw(self, v, dm) -> torch.Tensor:
        """PDEfunc(JACOB_FORW): Pr H^-1 (-T(U) Pext dm) with
        T(U) dm = dH (dm .* U) (discrete_helmholtz.m:287)."""
        v = self._model_tensor(v)
        dm = self._model_tensor(dm)
        nsrc = self.Q.shape[1]
        out = torch.zeros(self.model.nrec, len(self.model.freq) * nsrc, dtype=torch.complex128,
                          device=self.device)
        for k, freq in enumerate(self.model.freq):
            S = self._setup(v, freq)
            dwn_eff = self._dwn_times(S, freq, dm)
            for j in range(nsrc):
                U = self._solve(self._source(S, k, j), 1)
                rhs = self.H.apply_deriv(dwn_eff, U)
                dU = self._solve(-rhs, 1)
                out[:, k * nsrc + j] = self._sample(S, dU)
        return out

    def jacob_adj(self, v, dd) -> torch.Tensor:
        """PDEfunc(JACOB_ADJ): comp2phys sum Re(conj(U) .* dH^H H^-H(-Pr^T dd))."""
        v = self._model_tensor(v)
        dd = self._data(dd)
        nsrc = self.Q.shape[1]
        g = torch.zeros(int(np.prod(self.n)), dtype=torch.float64, device=self.device)
        for k, freq in enumerate(self.model.freq):
            S = self._setup(v, freq)
            gpml = self._grad_buffers(S)
            Us = self._solve_many([self._source(S, k, j) for j in range(nsrc)], 1)
            Vs = self._solve_many([self._inject_rec(S, dd[:, k * nsrc + j], -1.0) for j in range(nsrc)], 2)
            for U, V in zip(Us, Vs):
                self._grad_term(S, freq, U, V, g, gpml)
            self._grad_finish(S, g, gpml)
        return g

    def hess_gn(self, v, dm) -> torch.Tensor:
        """PDEfunc(HESS_GN): J^T J dm with the two extra solves of
        PDEfunc.m:437-441."""
        v = self._model_tensor(v)
        dm = self._model_tensor(dm)
        nsrc = self.Q.shape[1]
        g = torch.zeros(int(np.prod(self.n)), dtype=torch.float64, device=self.device)
        for k, freq in enumerate(self.model.freq):
            S = self._setup(v, freq)
            gpml = self._grad_buffers(S)
            dwn_eff = self._dwn_times(S, freq, dm)
            for j in range(nsrc):
                U = self._solve(self._source(S, k, j), 1)
                dU = self._solve(-self.H.apply_deriv(dwn_eff, U), 1)
                dU = self._solve(self._inject_rec(S, self._sample(S, dU), -1.0), 2)
                self._grad_term(S, freq, U, dU, g, gpml)
            self._grad_finish(S, g, gpml)
        return g


    def hess(self, v, dm, Dobs) -> torch.Tensor:
        r"""PDEfunc(HESS) with the LS misfit (d2phi = I), PDEfunc.m:444-450:
            dU = H \ (-T(U) dm),  V = H' \ (-Pr' dphi),
            dV = H' \ (-T(V) dm - Pr' Pr dU),
            out += comp2phys sum Re(DT_adj(U, dm, dU) V + T(U)' dV)
        with T(u) dm = dH (dm .* u) and DT_adj(U, dm, dU) V = conj(U) .* (dm .*
        ddH' V) + conj(dU) .* (dH' V) (discrete_helmholtz.m:286-293); dm is on
        the extended grid (phys2comp).  T(V) dm is the forward derivative
        operator applied to V, as the reference forms it."""
        v = self._model_tensor(v)
        dm = self._model_tensor(dm)
        D = self._data(Dobs)
        nsrc = self.Q.shape[1]
        g = torch.zeros(int(np.prod(self.n)), dtype=torch.float64, device=self.device)
        for k, freq in enumerate(self.model.freq):
            S = self._setup(v, freq)
            gpml = self._grad_buffers(S)
            dwn_eff = self._dwn_times(S, freq, dm)
            ddwn_dm = None
            if self.model.unit == "m/s":  # ddwn = 0 for the slowness units (param2wavenum.m:36-49)
                _, _, ddwn = self.H.param2wavenum(v, self.n, S["p6"], freq, self.model.unit, derivs=2)
                ddwn_dm = ddwn
                dm_pml = self.H.extend(dm, self.n, S["p6"])
            for j in range(nsrc):
                U = self._solve(self._source(S, k, j), 1)
                r = self._sample(S, U) - D[:, k * nsrc + j]
                dU = self._solve(-self.H.apply_deriv(dwn_eff, U), 1)
                V = self._solve(self._inject_rec(S, r, -1.0), 2)
                rhs = self._inject_rec(S, self._sample(S, dU), -1.0) - self.H.apply_deriv(dwn_eff, V)
                dV = self._solve(rhs, 2)
                self._grad_term(S, freq, dU, V, g, gpml)
                self._grad_term(S, freq, U, dV, g, gpml)
                if ddwn_dm is not None:
                    # conj(U) .* (dm .* ddH' V) = conj(dm .* U) .* (ddH' V) for real dm
                    Udm = U * dm_pml
                    if gpml is None:
                        self.H.grad_accum(Udm, V, g, dwn=ddwn_dm, grid="phys")
                    else:
                        self.H.grad_accum(Udm, V, gpml, dwn=ddwn_dm, grid="pml")
            self._grad_finish(S, g, gpml)
        return g


def PDEfunc(func: str, v, Q, inp, Dobs, model: FWIModel, opts: Optional[FWIOpts] = None, problem=None):
    """PDEfunc(func, v, Q, input, Dobs, model, params) for the 3D operto27
    scheme.  Returns (f, g) for OBJ, the nrec x (nsrc nfreq) data for
    FORW_MODEL / JACOB_FORW, and the physical-grid image for JACOB_ADJ /
    HESS_GN / HESS (float64 device tensors; HESS uses Dobs)."""
    if func not in MODES:
        raise ValueError("Must choose a preset function.")
    P = problem or FWIProblem(model, opts, Q=Q)
    if func == OBJ:
        return P.objective(v, Dobs, need_grad=True)
    if func == FORW_MODEL:
        return P.forw_model(v)
    if func == JACOB_FORW:
        return P.jacob_forw(v, inp)
    if func == JACOB_ADJ:
        return P.jacob_adj(v, inp)
    if func == HESS:
        return P.hess(v, inp, Dobs)
    return P.hess_gn(v, inp)


class FWIObjective(torch.autograd.Function):
    """f(m) = PDEfunc(OBJ) as a differentiable function of the model m
    (float64 tensor on the physical grid).

    Forward runs the forward solves and keeps, per (frequency, source), the
    receiver residual and -- when HBM allows -- the wavefield U.  Backward runs
    the adjoint solves and accumulates the gradient in the HIP gradient
    kernel; wavefields that were not kept are recomputed first.
    wavefields: "store", "recompute" or "auto" (store while the wavefields fit
    in half of the free device memory)."""

    @staticmethod
    def forward(ctx, m: torch.Tensor, problem: FWIProblem, Dobs, wavefields: str = "auto"):
        P = problem
        v = P._model_tensor(m)
        D = P._data(Dobs)
        nsrc = P.Q.shape[1]
        npairs = len(P.model.freq) * nsrc
        if wavefields not in ("auto", "store", "recompute"):
            raise ValueError("wavefields must be 'auto', 'store' or 'recompute'")
        keep = ctx.needs_input_grad[0] and wavefields != "recompute"
        if keep and wavefields == "auto":
            free, _ = torch.cuda.mem_get_info(P.device)
            keep = npairs * P.pair_bytes(v) <= free // 2
        f = 0.0
        Us, R = [], []
        for k, freq in enumerate(P.model.freq):
            S = P._setup(v, freq)
            Uk = P._solve_many([P._source(S, k, j) for j in range(nsrc)], 1)
            for j, U in enumerate(Uk):
                r = P._sample(S, U) - D[:, k * nsrc + j]
                f += 0.5 * float(torch.vdot(r, r).real)
                Us.append(U if keep else None)
                R.append(r)
            del Uk
        ctx.problem = P
        ctx.Us = Us
        ctx.R = R
        ctx.save_for_backward(v)
        ctx.m_shape = m.shape
        ctx.m_dtype = m.dtype
        ctx.stored = keep
        return torch.tensor(f, dtype=torch.float64, device=P.device)

    @staticmethod
    def backward(ctx, grad_out):
        P = ctx.problem
        (v,) = ctx.saved_tensors
        nsrc = P.Q.shape[1]
        g = torch.zeros(int(np.prod(P.n)), dtype=torch.float64, device=P.device)
        for k, freq in enumerate(P.model.freq):
            S = P._setup(v, freq)
            gpml = P._grad_buffers(S)
            # sources in chunks of the batch size: wavefields that were not
            # kept are recomputed one chunk at a time
            step = max(1, P.opts.batch)
            for j0 in range(0, nsrc, step):
                idx = [k * nsrc + j for j in range(j0, min(nsrc, j0 + step))]
                missing = [i for i in idx if ctx.Us[i] is None]
                if missing:
                    Ur = P._solve_many([P._source(S, k, i - k * nsrc) for i in missing], 1)
                    for i, U in zip(missing, Ur):
                        ctx.Us[i] = U
                Vs = P._solve_many([P._inject_rec(S, ctx.R[i], -1.0) for i in idx], 2)
                for i, V in zip(idx, Vs):
                    P._grad_term(S, freq, ctx.Us[i], V, g, gpml)
                    ctx.Us[i] = None
            P._grad_finish(S, g, gpml)
        P._zero_boundary(g)
        return (g * grad_out).reshape(ctx.m_shape).to(ctx.m_dtype), None, None, None


def fwi_misfit(m: torch.Tensor, problem: FWIProblem, Dobs, wavefields: str = "auto") -> torch.Tensor:
    """Differentiable least-squares misfit of the model m (see FWIObjective)."""
    return FWIObjective.apply(m, problem, Dobs, wavefields)
