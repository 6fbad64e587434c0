"""Multigrid-preconditioned FGMRES (the ML_GMRES preconditioner of the
reference: tools/solvers/Multigrid/ML_GMRES.m:34-72, opMultigrid.m:71-110,
construct_multigrid.m:10-63, GMG_V.m) over the C ABI.

The level models are built as construct_multigrid.m does: level l coarsens
the physical model of level l-1 to ceil(n/2) points per axis with the cubic
interpolation of fine2coarse.m (opLInterp1D over linspace(0,1,n)), doubles
the spacing, uses ceil(npml / 2^l) PML points, and discretises it again
(discrete_helmholtz, on the device).  The V-cycle and every solve run in the
HIP library (helm3d_solve_mg).
"""
from __future__ import annotations

import ctypes
import math
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .core import Helm3d, SolveInfo, _stream_ptr
from .fwi import lagrange_cubic


def coarsen_model(v, n, n_coarse):
    """to_coarse * vec(v) of fine2coarse(n, d, 2d) with cubic interpolation
    (fine2coarse.m:28-53): separable, x fastest.  v: numpy array (host) or
    torch tensor (the product runs it on the model's device)."""
    if isinstance(v, torch.Tensor):
        V = v.reshape(n[2], n[1], n[0]).to(torch.float64)
        for ax, a in ((2, 0), (1, 1), (0, 2)):
            W = torch.from_numpy(lagrange_cubic(np.linspace(0, 1, n[a]), np.linspace(0, 1, n_coarse[a])))
            V = torch.movedim(torch.tensordot(W.to(V.device), V, dims=([1], [ax])), 0, ax)
        return V.reshape(-1).contiguous()
    V = np.asarray(v, dtype=np.float64).reshape(n[2], n[1], n[0])
    for ax, a in ((2, 0), (1, 1), (0, 2)):  # numpy axis <-> grid axis a
        W = lagrange_cubic(np.linspace(0, 1, n[a]), np.linspace(0, 1, n_coarse[a]))
        V = np.moveaxis(np.tensordot(W, V, axes=([1], [ax])), 0, ax)
    return V.reshape(-1)


class MLGMRES:
    """Multigrid hierarchy for one model grid, and its solver.

    v: physical model (prod(n_phys) values, x fastest) in ``unit``, a numpy
    array or a device tensor; npml: PML points of level 0 on every face
    (construct_multigrid.m:21 takes comp_grid.pml(1,1)); free_surface drops
    the top (z lo) PML on every level, as discrete_helmholtz does;
    k = (ks_outer, ks_inner, kc_outer, kc_inner) of ML_GMRES.m:3-13.
    autotune: time the stencil launch shapes of every level once (the FP64
    shape for the level-0 operator, the complex64 shape for the smoothers and
    coarse levels); shapes survive set_model and never change results."""

    def __init__(self, v, n_phys, h, npml: int, freq: float, unit: str = "m/s", nlevels: int = 3,
                 k: Sequence[int] = (3, 5, 3, 5), device=None, free_surface: bool = False,
                 autotune: bool = True):
        device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.device = device
        self.k = tuple(int(a) for a in k)
        self.unit = unit
        n = tuple(int(a) for a in n_phys)
        h = tuple(float(a) for a in np.asarray(h, dtype=np.float64).reshape(-1)) if np.size(h) == 3 else (float(h),) * 3
        self.ops, self.smoothers, self.grids = [], [], []
        for lvl in range(nlevels):
            if lvl > 0:
                n = tuple(int(math.ceil(a / 2)) for a in n)
                h = tuple(2.0 * a for a in h)
            p = int(math.ceil(int(npml) / 2 ** lvl))
            p6 = (p, p, p, p, 0 if free_surface else p, p)
            nt = tuple(n[a] + p6[2 * a] + p6[2 * a + 1] for a in range(3))
            self.grids.append(dict(n_phys=n, nt=nt, h=h, npml=p, p6=p6))
            self.ops.append(Helm3d(device=device))
            self.smoothers.append(Helm3d(device=device))
        self.set_model(v, freq)
        if autotune:
            for lvl in range(nlevels):
                self.ops[lvl].autotune(mode=1, precision="c128" if lvl == 0 else "c64", reps=5)
                self.smoothers[lvl].autotune(mode=1, precision="c64", reps=5)
        self._mg = ctypes.c_void_p()
        L = len(self.ops)
        a_arr = (ctypes.c_void_p * L)(*[H._ctx.value for H in self.ops])
        s_arr = (ctypes.c_void_p * L)(*[H._ctx.value for H in self.smoothers])
        with torch.cuda.device(device):
            check(lib().helm3d_set_stream(self.ops[0]._ctx, ctypes.c_void_p(_stream_ptr())))
            check(lib().helm3d_mg_create(L, a_arr, s_arr, ctypes.byref(self._mg)))

    def set_model(self, v, freq: float):
        """Load the model of every level (construct_multigrid.m:46-49: cubic
        coarsening of the previous level's physical model, discrete_helmholtz
        on the coarse grid); grids and transfers stay as built."""
        vl = v if isinstance(v, torch.Tensor) else torch.from_numpy(np.asarray(v, dtype=np.float64).reshape(-1))
        vl = vl.reshape(-1).to(device=self.device, dtype=torch.float64)
        for lvl, g in enumerate(self.grids):
            if lvl > 0:
                vl = coarsen_model(vl, self.grids[lvl - 1]["n_phys"], g["n_phys"])
            A, S = self.ops[lvl], self.smoothers[lvl]
            wn, _, _ = A.param2wavenum(vl, g["n_phys"], g["p6"], freq, self.unit, derivs=0)
            A.set_model(g["nt"], g["h"], g["p6"], wn)
            S.set_model(g["nt"], g["h"], g["p6"], wn)
        self.freq = float(freq)

    @property
    def H(self) -> Helm3d:
        """The level-0 operator."""
        return self.ops[0]

    def close(self):
        if getattr(self, "_mg", None) and self._mg.value:
            lib().helm3d_mg_destroy(self._mg)
            self._mg = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _bind(self):
        """Run the hierarchy on the caller's current stream (where b, x and
        the outputs were allocated or zeroed)."""
        check(lib().helm3d_mg_set_stream(self._mg, ctypes.c_void_p(_stream_ptr())))

    def transfer(self, level: int, x: torch.Tensor, direction: str) -> torch.Tensor:
        """"prolong": P x (level+1 -> level); "restrict": P^T x / 8 (level -> level+1)."""
        if direction == "prolong":
            out = torch.zeros(int(np.prod(self.grids[level]["nt"])), dtype=torch.complex128, device=self.device)
            d = 0
        else:
            out = torch.empty(int(np.prod(self.grids[level + 1]["nt"])), dtype=torch.complex128, device=self.device)
            d = 1
        x = x.to(device=self.device, dtype=torch.complex128).contiguous()
        with torch.cuda.device(self.device):
            self._bind()
            check(lib().helm3d_mg_transfer(self._mg, int(level), d, ctypes.c_void_p(x.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr())))
        return out

    def solve(self, b: torch.Tensor, mode: int = 1, tol: float = 1e-6, maxinnerit: int = 20, maxit: int = 10000,
              x0: Optional[torch.Tensor] = None, raise_on_notconv: bool = False, precision: str = "c128"):
        """FGMRES on the level-0 operator preconditioned by the V-cycle;
        precision "mixed" runs the V-cycle in complex64 (the outer FGMRES and
        its residuals stay FP64)."""
        if precision not in ("c128", "mixed"):
            raise ValueError("precision must be 'c128' or 'mixed'")
        H = self.ops[0]
        b = H._vec(b, torch.complex128)
        x = torch.zeros_like(b) if x0 is None else H._vec(x0, torch.complex128).clone()
        o = Helm3d._opts(tol, maxinnerit, maxit, x0 is not None)
        info = _lib.SolveInfo()
        kk = (ctypes.c_int * 4)(*self.k)
        with torch.cuda.device(self.device):
            self._bind()
            fn = lib().helm3d_solve_mg if precision == "c128" else lib().helm3d_solve_mg_mixed
            code = fn(self._mg, int(mode), ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                      ctypes.byref(o), kk, ctypes.byref(info))
        check(code, allow=() if raise_on_notconv else (_lib.HELM3D_ERR_NOTCONV,))
        return x, SolveInfo(int(info.iters), int(info.cycles), float(info.relres), float(info.quasi_relres),
                            bool(info.converged))


class MGPool:
    """Several multigrid hierarchies of one model, each on its own stream and
    host thread, so that the solves of many right-hand sides overlap on the
    device (the coarse levels of one V-cycle leave most of the GPU idle).
    Arguments as for MLGMRES, plus ``workers`` (hierarchies)."""

    def __init__(self, v, n_phys, h, npml: int, freq: float, unit: str = "m/s", nlevels: int = 3,
                 k: Sequence[int] = (3, 5, 3, 5), device=None, free_surface: bool = False,
                 autotune: bool = True, workers: int = 2):
        device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.device = device
        self.streams = [torch.cuda.Stream(device) for _ in range(max(1, int(workers)))]
        self.hier = []
        main = torch.cuda.current_stream(device)
        for s in self.streams:
            s.wait_stream(main)  # v may still be in flight on the caller's stream
            with torch.cuda.device(device), torch.cuda.stream(s):
                self.hier.append(MLGMRES(v, n_phys, h, npml, freq, unit, nlevels, k, device, free_surface,
                                         autotune))
        for s in self.streams:
            s.synchronize()

    @property
    def H(self) -> Helm3d:
        return self.hier[0].H

    @property
    def grids(self):
        return self.hier[0].grids

    def set_model(self, v, freq: float):
        main = torch.cuda.current_stream(self.device)
        for s, M in zip(self.streams, self.hier):
            s.wait_stream(main)  # v may still be in flight on the caller's stream
            with torch.cuda.device(self.device), torch.cuda.stream(s):
                M.set_model(v, freq)
        for s in self.streams:
            s.synchronize()

    def close(self):
        for M in self.hier:
            M.close()
        self.hier = []

    def solve_many(self, bs: Sequence[torch.Tensor], mode: int = 1, tol: float = 1e-6, maxinnerit: int = 20,
                   maxit: int = 10000, precision: str = "c128"):
        """x_i = A^-1 b_i (A = H, or H^H when mode != 1) for every b_i, the
        right-hand sides taken in turn by the hierarchies.  Returns
        (xs, infos) in the order of bs; every x is ready on the caller's
        current stream."""
        import threading
        main = torch.cuda.current_stream(self.device)
        ready = torch.cuda.Event()
        ready.record(main)
        out: list = [None] * len(bs)
        lock = threading.Lock()
        nxt = [0]
        errors: list = []

        def work(w: int):
            s, M = self.streams[w], self.hier[w]
            try:
                with torch.cuda.device(self.device), torch.cuda.stream(s):
                    s.wait_event(ready)
                    while True:
                        with lock:
                            i = nxt[0]
                            nxt[0] += 1
                        if i >= len(bs) or errors:
                            break
                        out[i] = M.solve(bs[i], mode=mode, tol=tol, maxinnerit=maxinnerit, maxit=maxit,
                                         precision=precision)
            except BaseException as e:  # re-raised on the calling thread
                errors.append(e)

        threads = [threading.Thread(target=work, args=(w,)) for w in range(len(self.hier))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for s in self.streams:
            done = torch.cuda.Event()
            done.record(s)
            main.wait_event(done)
        if errors:
            raise errors[0]
        for x, _ in out:
            x.record_stream(main)
        return [x for x, _ in out], [inf for _, inf in out]
