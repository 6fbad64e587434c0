"""helm3d_apply_noalloc (the stencil with its scratch in a static device array
of the library, no device allocation) against helm3d_apply on a context that
holds the same model: two calls in a row, the second on a smaller grid, for
the forward and adjoint operators with real and complex wn.  Bar: 1e-12
relative (the PML tables are evaluated with the device cos instead of the
host's, which moves a last bit at most).  The axis limit is an argument
error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRIDS = [((40, 36, 30), (5, 6, 4, 5, 0, 6)), ((21, 18, 16), (3, 3, 2, 4, 3, 3))]  # second one smaller


def _noalloc(variant, n, h, npml, wn, wnc, x):
    import torch
    from helm3d._lib import check, lib
    y = torch.empty_like(x)
    ni = (ctypes.c_int * 3)(*n)
    hd = (ctypes.c_double * 3)(*h)
    pi = (ctypes.c_int * 6)(*npml)
    code = lib().helm3d_apply_noalloc(int(variant), ni, hd, pi, ctypes.c_void_p(wn.data_ptr()), int(wnc),
                                      ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    check(code)
    return y


@pytest.mark.parametrize("wnc", [False, True])
def test_noalloc_matches_context_apply(wnc):
    import torch
    from helm3d import Helm3d
    from helm3d._lib import HELM3D_ADJ, HELM3D_FORW
    dev = torch.device("cuda:0")
    h = (10.0, 11.0, 9.5)
    ys = []
    for step, (n, npml) in enumerate(GRIDS):
        print(f"noalloc call {step}: n={n}, npml={npml}, wnc={wnc}", flush=True)
        rng = np.random.default_rng(100 + step)
        N = int(np.prod(n))
        wn_np = rng.uniform(2e-4, 2e-3, N)
        if wnc:
            wn_np = wn_np + 1j * rng.uniform(-2e-4, 2e-4, N)
        wn = torch.from_numpy(wn_np).to(dev)
        x = torch.from_numpy(rng.standard_normal(N) + 1j * rng.standard_normal(N)).to(dev)
        H = Helm3d(n, h, np.array(npml).reshape(3, 2).T, wn)
        for variant, adjoint in ((HELM3D_FORW, False), (HELM3D_ADJ, True)):
            y0 = _noalloc(variant, n, h, npml, wn, wnc, x)
            ref = H.apply(x, adjoint=adjoint)
            torch.cuda.synchronize()
            err = float((y0 - ref).norm() / ref.norm())
            print(f"  variant {variant}: rel diff vs helm3d_apply {err:.2e}", flush=True)
            assert err <= 1e-12
            ys.append(y0)
        H.close()
    # the first call's outputs are not touched by the second call
    assert all(torch.isfinite(y).all() for y in ys)


def test_noalloc_rejects_oversized_axis():
    import torch
    from helm3d._lib import HELM3D_ERR_ARG, HELM3D_FORW, HELM3D_NOALLOC_MAX_AXIS, lib
    n = (HELM3D_NOALLOC_MAX_AXIS + 1, 2, 2)
    wn = torch.zeros(4, dtype=torch.float64, device="cuda:0")
    x = torch.zeros(4, dtype=torch.complex128, device="cuda:0")
    code = lib().helm3d_apply_noalloc(HELM3D_FORW, (ctypes.c_int * 3)(*n), (ctypes.c_double * 3)(1.0, 1.0, 1.0),
                                      (ctypes.c_int * 6)(0, 0, 0, 0, 0, 0), ctypes.c_void_p(wn.data_ptr()), 0,
                                      ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(x.data_ptr()), None)
    assert code == HELM3D_ERR_ARG
