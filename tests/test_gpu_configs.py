"""The benchmark configurations at their stated sizes, each checked with the
CPU oracle.

  C1  64^3 + 10-pt PML (84^3), 5 Hz, one impulse at (32,32,5): the plain
      complex128 FGMRES(20) solve to 1e-6 on the device against the oracle's
      FGMRES.m restatement driving the reference stencil (oracle/_ref when
      built, else the oracle's own): iteration count within 2 % + 2, the
      quasi-residual history of the first 40 steps to 1e-6, and the device
      solution's residual measured with the oracle operator in FP64.
  C3  the 240x240x140 benchmark model with 8 surface sources, batched mixed
      precision (FP64 flexible GMRES, complex64 BiCGStab inner iterations):
      every FP64 residual below 1e-6 with the oracle operator, and each
      solution within 1e-3 of a complex128 solve to 1e-10 (the error is
      printed per solve).  The single-right-hand-side mixed solve with an
      inner complex64 GMRES (the inner FGMRES of the configuration) is run on
      two of the sources at the same size.
  C4  the uneven batched workload at full size (8 grids between 96^3 and
      256x256x128 + 20-pt PML, 3-8 Hz, 130 impulses), split into batches
      that fit the device: all 130 residuals below 1e-6 with the oracle
      operator.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _wn(c):
    from helm3d.models import extend, vec, wavenumber
    return vec(wavenumber(extend(c.v, c.npml), c.freq))


def _relres(c, wn, b, x, variant="forw"):
    r = b - O.hmvp(variant, wn, (c.h,) * 3, c.nt, np.array(c.npml6), x)
    return float(np.linalg.norm(r) / np.linalg.norm(b))


@pytest.mark.timeout(600)
def test_c1_solve_matches_oracle():
    import torch
    from helm3d import Helm3d
    from helm3d.models import config_c1, point_source
    c = config_c1()
    wn = _wn(c)
    b = point_source(c, c.sources[0])
    print(f"C1: nt={c.nt}, {c.freq} Hz, source {c.sources[0]}; oracle FGMRES(20) on the CPU", flush=True)
    xo, io = O.fgmres("forw", wn, (c.h,) * 3, c.nt, np.array(c.npml6), b, tol=1e-6, maxinnerit=20,
                      use_ref=O.ref_available())
    dev = torch.device("cuda:0")
    H = Helm3d(c.nt, (c.h,) * 3, c.npml6, torch.from_numpy(wn).to(dev))
    x, info = H.solve(torch.from_numpy(b).to(dev), tol=1e-6, maxinnerit=20, return_hst=True)
    x = x.cpu().numpy()
    rel = _relres(c, wn, b, x)
    print(f"  device {info.iters} iterations (relres {rel:.3e}), oracle {io['iters']} (relres {io['relres']:.3e})",
          flush=True)
    assert info.converged and rel < 1e-6 and abs(rel - info.relres) < 1e-9
    assert abs(info.iters - io["iters"]) <= 0.02 * io["iters"] + 2
    k = min(40, len(info.hst), len(io["hst"]))
    assert np.allclose(info.hst[:k], io["hst"][:k], rtol=1e-6, atol=0)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-3


@pytest.mark.timeout(600)
def test_c3_batched_mixed_full_size():
    import torch
    from helm3d import BatchSolver, Helm3d
    from helm3d.mg import MLGMRES
    from helm3d.models import config_c3, point_source, vec
    c = config_c3()
    wn = _wn(c)
    dev = torch.device("cuda:0")
    H = Helm3d(c.nt, (c.h,) * 3, c.npml6, torch.from_numpy(wn).to(dev))
    bh = [point_source(c, s) for s in c.sources]
    bs = [torch.from_numpy(b).to(dev) for b in bh]
    print(f"C3: nt={c.nt}, {len(bs)} sources, batched mixed precision", flush=True)
    S = BatchSolver(dev)
    xs, infos = S.solve([H] * len(bs), bs, method="mixed", tol=1e-6, maxinnerit=8, inner_maxit=20)
    M = MLGMRES(vec(c.v), c.n_phys, c.h, c.npml, c.freq, nlevels=4, device=dev, autotune=False)
    try:
        for k, (b, x, inf) in enumerate(zip(bh, xs, infos)):
            rel = _relres(c, wn, b, x.cpu().numpy())
            xr, ir = M.solve(bs[k], tol=1e-10, precision="c128")
            assert ir.converged
            err = float((x - xr).norm() / xr.norm())
            print(f"  source {c.sources[k]}: {inf.iters} outer steps, relres {rel:.3e}, error vs complex128 {err:.2e}",
                  flush=True)
            assert inf.converged and rel < 1e-6 and abs(rel - inf.relres) < 1e-9
            assert err < 1e-3
        # the inner-GMRES form of the mixed solve (single right-hand side)
        for k in (0, len(bs) - 1):
            xg, ig = H.solve_mixed(bs[k], tol=1e-6, inner="gmres", inner_tol=1e-2)
            rel = _relres(c, wn, bh[k], xg.cpu().numpy())
            xr, _ = M.solve(bs[k], tol=1e-10, precision="c128")
            err = float((xg - xr).norm() / xr.norm())
            print(f"  inner GMRES, source {c.sources[k]}: {ig.iters} outer / {ig.inner_iters} inner, "
                  f"relres {rel:.3e}, error vs complex128 {err:.2e}", flush=True)
            assert ig.converged and rel < 1e-6 and err < 1e-3
    finally:
        M.close()


@pytest.mark.timeout(900)
def test_c4_uneven_workload_full_size():
    import torch
    from helm3d import BatchSolver, Helm3d
    from helm3d.models import config_c4, point_source
    subs = config_c4()
    dev = torch.device("cuda:0")
    ops, bs, meta = [], [], []
    for c in subs:
        wn = _wn(c)
        Hk = Helm3d(c.nt, (c.h,) * 3, c.npml6, torch.from_numpy(wn).to(dev))
        for s in c.sources:
            ops.append(Hk)
            bs.append(torch.from_numpy(point_source(c, s)).to(dev))
            meta.append((c, wn))
    print(f"C4: {len(subs)} grids {[c.nt for c in subs]}, {len(bs)} right-hand sides", flush=True)
    S = BatchSolver(dev)
    xs, infos, chunks = S.solve_chunked(ops, bs, method="mixed", tol=1e-6, maxinnerit=8, inner_maxit=20)
    print(f"  batches {[len(ch) for ch in chunks]}", flush=True)
    assert sorted(i for ch in chunks for i in ch) == list(range(len(bs)))
    worst = 0.0
    for i, ((c, wn), b, x, inf) in enumerate(zip(meta, bs, xs, infos)):
        rel = _relres(c, wn, b.cpu().numpy(), x.cpu().numpy())
        worst = max(worst, rel)
        assert inf.converged and rel < 1e-6 and abs(rel - inf.relres) < 1e-9, (i, c.name, rel, inf)
    print(f"  all {len(bs)} residuals below 1e-6 with the oracle operator (worst {worst:.3e}); outer steps "
          f"{min(i.iters for i in infos)}-{max(i.iters for i in infos)}", flush=True)
