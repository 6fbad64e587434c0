"""Batched solves over the C ABI: many right-hand sides on one or several
operators (grids, frequencies, models) advanced together, one launch per
kernel type and Krylov step for the whole batch (``helm3d_batch_solve``).

``method="plain"`` is the multi-right-hand-side GMRES of
tools/solvers/Krylov/FGMRES_MT.m:40-135 in complex128 (lockstep cycles,
max-over-items stopping); ``method="mixed"`` is FP64 flexible GMRES with a
fixed number of complex64 BiCGStab iterations as preconditioner, items
stopping individually.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import torch

from . import _lib
from ._lib import check, lib
from .core import Helm3d, SolveInfo, _stream_ptr


class BatchSolver:
    """Owns the batch workspace (Krylov bases and inner vectors of every item)
    on one device; reusable across calls."""

    def __init__(self, device=None):
        device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.device = device
        self._b = ctypes.c_void_p()
        with torch.cuda.device(device):
            check(lib().helm3d_batch_create(device.index if device.index is not None else -1,
                                            ctypes.c_void_p(_stream_ptr()), ctypes.byref(self._b)))

    def close(self):
        if getattr(self, "_b", None) and self._b.value:
            lib().helm3d_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self, ops: Sequence[Helm3d], bs: Sequence[torch.Tensor], mode: int = 1, method: str = "mixed",
                 inner_maxit: int = 25, tile_rows=(4, 8), prefetch=(1, 2), divisors=(1, 2, 3, 4),
                 rows_per_thread=(1, 2)) -> dict:
        """Choose the stencil launch shape of a batch by timing a short batched
        solve (one restart cycle of two outer steps) for every candidate: 64 x
        tile_rows column tiles (64 x 8 with one or two rows per thread),
        prefetch depth, and z chunks of nz/divisor for every operator.  The
        fastest shape is left set on the operators; later solves converge to
        the same tolerance whatever the shape (the fused BiCGStab partial sums
        follow the shape, so iteration counts can move by a few)."""
        uniq = list({id(H): H for H in ops}.values())
        best = None
        for ty in tile_rows:
            for ry in (rows_per_thread if ty == 8 else (1,)):
                for pf in prefetch:
                    for d in divisors:
                        for H in uniq:
                            H.set_launch((H.n[2] + d - 1) // d, ty, pf, ry)
                        torch.cuda.synchronize(self.device)
                        t = []
                        for _ in range(2):
                            e0 = torch.cuda.Event(enable_timing=True)
                            e1 = torch.cuda.Event(enable_timing=True)
                            e0.record()
                            self.solve(ops, bs, mode=mode, method=method, tol=1e-30, maxinnerit=2, maxit=1,
                                       inner_maxit=inner_maxit)
                            e1.record()
                            torch.cuda.synchronize(self.device)
                            t.append(e0.elapsed_time(e1))
                        cand = (min(t), ty, ry, pf, d)
                        if best is None or cand < best:
                            best = cand
        _, ty, ry, pf, d = best
        for H in uniq:
            H.set_launch((H.n[2] + d - 1) // d, ty, pf, ry)
        return dict(tile_rows=ty, rows_per_thread=ry, prefetch=pf, divisor=d, ms=best[0])

    @staticmethod
    def item_bytes(H: Helm3d, method: str = "mixed", maxinnerit: int = 20) -> int:
        """Device workspace one right-hand side on H adds to a batch."""
        meth = {"plain": _lib.HELM3D_BATCH_PLAIN, "mixed": _lib.HELM3D_BATCH_MIXED}[method]
        out = ctypes.c_size_t()
        check(lib().helm3d_batch_item_bytes(H._ctx, meth, int(maxinnerit), ctypes.byref(out)))
        return int(out.value)

    def plan(self, ops: Sequence[Helm3d], method: str = "mixed", maxinnerit: int = 20,
             budget_bytes: Optional[int] = None) -> List[List[int]]:
        """Split the items, in the given order, into consecutive batches whose
        workspace fits ``budget_bytes`` (default: 85 % of the device memory
        free now).  Items of one operator should be adjacent: a mixed batch
        lasts as long as its slowest item."""
        if budget_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            budget_bytes = int(0.85 * free)
        cost = {}
        chunks: List[List[int]] = []
        used = 0
        for i, H in enumerate(ops):
            if id(H) not in cost:
                cost[id(H)] = self.item_bytes(H, method, maxinnerit)
            c = cost[id(H)]
            if c > budget_bytes:
                raise MemoryError(f"one right-hand side needs {c} B of workspace, budget {budget_bytes} B")
            if not chunks or used + c > budget_bytes or len(chunks[-1]) >= _lib.HELM3D_BATCH_MAX_ITEMS:
                chunks.append([])
                used = 0
            chunks[-1].append(i)
            used += c
        return chunks

    def solve_chunked(self, ops: Sequence[Helm3d], bs: Sequence[torch.Tensor], mode: int = 1,
                      method: str = "mixed", budget_bytes: Optional[int] = None, **kw
                      ) -> Tuple[List[torch.Tensor], List[SolveInfo], List[List[int]]]:
        """solve() over any number of right-hand sides: the items are split
        by plan() into batches that fit the device, solved one batch after
        the other.  Returns (xs, infos, batches)."""
        maxinnerit = kw.get("maxinnerit", 20)
        chunks = self.plan(ops, method, maxinnerit, budget_bytes)
        xs: List[Optional[torch.Tensor]] = [None] * len(ops)
        infos: List[Optional[SolveInfo]] = [None] * len(ops)
        for ch in chunks:
            x, inf = self.solve([ops[i] for i in ch], [bs[i] for i in ch], mode=mode, method=method, **kw)
            for j, i in enumerate(ch):
                xs[i], infos[i] = x[j], inf[j]
        return xs, infos, chunks

    def solve(self, ops: Sequence[Helm3d], bs: Sequence[torch.Tensor], mode: int = 1, method: str = "mixed",
              tol: float = 1e-6, maxinnerit: int = 20, maxit: int = 10000, inner_maxit: int = 25,
              raise_on_notconv: bool = False) -> Tuple[List[torch.Tensor], List[SolveInfo]]:
        """x_i = A_i^-1 b_i for every item (A_i = ops[i] or its adjoint when
        mode != 1).  ops may repeat the same operator for several
        right-hand sides."""
        if len(ops) != len(bs) or not ops:
            raise ValueError("need one operator per right-hand side")
        meth = {"plain": _lib.HELM3D_BATCH_PLAIN, "mixed": _lib.HELM3D_BATCH_MIXED}[method]
        bv = [H._vec(b, torch.complex128) for H, b in zip(ops, bs)]
        xs = [torch.empty_like(b) for b in bv]
        n = len(ops)
        ctxs = (ctypes.c_void_p * n)(*[H._ctx.value for H in ops])
        bp = (ctypes.c_void_p * n)(*[b.data_ptr() for b in bv])
        xp = (ctypes.c_void_p * n)(*[x.data_ptr() for x in xs])
        o = Helm3d._opts(tol, maxinnerit, maxit, False, inner_tol=0.0, inner_maxit=inner_maxit)
        infos = (_lib.SolveInfo * n)()
        with torch.cuda.device(self.device):
            check(lib().helm3d_batch_set_stream(self._b, ctypes.c_void_p(_stream_ptr())))
            code = lib().helm3d_batch_solve(self._b, n, ctxs, int(mode), bp, xp, meth, ctypes.byref(o), infos)
        check(code, allow=() if raise_on_notconv else (_lib.HELM3D_ERR_NOTCONV,))
        out = [SolveInfo(int(i.iters), int(i.cycles), float(i.relres), float(i.quasi_relres), bool(i.converged),
                         inner_iters=int(i.inner_iters)) for i in infos]
        return xs, out
