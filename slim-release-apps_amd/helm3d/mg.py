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


def coarsen_model(v: np.ndarray, n, n_coarse) -> np.ndarray:
    """to_coarse * vec(v) of fine2coarse(n, d, 2d) with cubic interpolation
    (fine2coarse.m:28-53): separable, x fastest."""
    V = np.asarray(v, dtype=np.float64).reshape(n[2], n[1], n[0])
    for ax, a in ((2, 0), (1, 1), (0, 2)):  # numpy axis <-> grid axis a
        W = lagrange_cubic(np.linspace(0, 1, n[a]), np.linspace(0, 1, n_coarse[a]))
        V = np.moveaxis(np.tensordot(W, V, axes=([1], [ax])), 0, ax)
    return V.reshape(-1)


class MLGMRES:
    """Multigrid hierarchy for one model and frequency, and its solver.

    v: physical model (prod(n_phys) values, x fastest) in ``unit``;
    k = (ks_outer, ks_inner, kc_outer, kc_inner) of ML_GMRES.m:3-13."""

    def __init__(self, v, n_phys, h, npml: int, freq: float, unit: str = "m/s", nlevels: int = 3,
                 k: Sequence[int] = (3, 5, 3, 5), device=None):
        device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.device = device
        self.k = tuple(int(a) for a in k)
        n = tuple(int(a) for a in n_phys)
        h = tuple(float(a) for a in np.asarray(h, dtype=np.float64).reshape(-1)) if np.size(h) == 3 else (float(h),) * 3
        vl = np.asarray(v, dtype=np.float64).reshape(-1)
        self.ops, self.smoothers, self.grids = [], [], []
        for lvl in range(nlevels):
            if lvl > 0:
                nc = tuple(int(math.ceil(a / 2)) for a in n)
                vl = coarsen_model(vl, n, nc)
                n = nc
                h = tuple(2.0 * a for a in h)
            p = int(math.ceil(npml / 2 ** lvl))
            p6 = (p,) * 6
            A = Helm3d(device=device)
            wn, _, _ = A.param2wavenum(torch.from_numpy(vl).to(device), n, p6, freq, unit, derivs=0)
            nt = tuple(a + 2 * p for a in n)
            A.set_model(nt, h, p6, wn)
            S = Helm3d(nt, h, p6, wn)
            self.ops.append(A)
            self.smoothers.append(S)
            self.grids.append(dict(n_phys=n, nt=nt, h=h, npml=p))
        self._mg = ctypes.c_void_p()
        L = len(self.ops)
        a_arr = (ctypes.c_void_p * L)(*[H._ctx.value for H in self.ops])
        s_arr = (ctypes.c_void_p * L)(*[H._ctx.value for H in self.smoothers])
        with torch.cuda.device(device):
            check(lib().helm3d_set_stream(self.ops[0]._ctx, ctypes.c_void_p(_stream_ptr())))
            check(lib().helm3d_mg_create(L, a_arr, s_arr, ctypes.byref(self._mg)))

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
            check(lib().helm3d_mg_transfer(self._mg, int(level), d, ctypes.c_void_p(x.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr())))
        return out

    def solve(self, b: torch.Tensor, mode: int = 1, tol: float = 1e-6, maxinnerit: int = 20, maxit: int = 10000,
              x0: Optional[torch.Tensor] = None, raise_on_notconv: bool = False):
        """FGMRES on the level-0 operator preconditioned by the V-cycle."""
        H = self.ops[0]
        b = H._vec(b, torch.complex128)
        x = torch.zeros_like(b) if x0 is None else H._vec(x0, torch.complex128).clone()
        o = Helm3d._opts(tol, maxinnerit, maxit, x0 is not None)
        info = _lib.SolveInfo()
        kk = (ctypes.c_int * 4)(*self.k)
        with torch.cuda.device(self.device):
            code = lib().helm3d_solve_mg(self._mg, int(mode), ctypes.c_void_p(b.data_ptr()),
                                         ctypes.c_void_p(x.data_ptr()), ctypes.byref(o), kk, ctypes.byref(info))
        check(code, allow=() if raise_on_notconv else (_lib.HELM3D_ERR_NOTCONV,))
        return x, SolveInfo(int(info.iters), int(info.cycles), float(info.relres), float(info.quasi_relres),
                            bool(info.converged))
