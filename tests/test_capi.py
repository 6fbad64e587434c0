"""CPU checks of the C ABI library: it loads and exports every symbol that
include/helm3d.h declares (no compute without a GPU)."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "helm3d.h")
LIB = os.path.join(REPO, "slim-release-apps_amd", "helm3d", "libhelm3d.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(helm3d_\w+)\s*\(", text)))


def test_header_declares_the_drop_in_family():
    names = declared()
    for v in ("forw", "adj", "forw_deriv", "adj_deriv", "forw_wnc", "adj_wnc"):
        assert f"helm3d_do_Hmvp_{v}" in names
    for n in ("helm3d_mvp", "helm3d_create", "helm3d_set_model", "helm3d_apply", "helm3d_solve",
              "helm3d_solve_mixed", "helm3d_fwi_grad_accum"):
        assert n in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libhelm3d.so not built: run __graft_entry__.build()")
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    import helm3d
    assert set(declared()) <= set(helm3d.EXPORTED)


def test_error_paths_without_gpu():
    """Size checks happen before any device call (MEX Nsize/Xsize ids)."""
    import numpy as np
    import helm3d
    from helm3d._lib import HELM3D_ERR_NSIZE, HELM3D_ERR_XSIZE
    L = helm3d.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    z = np.zeros(8)
    P = lambda a: a.ctypes.data_as(dp)  # noqa: E731
    n2 = np.array([2.0, 2.0])
    code = L.helm3d_mvp(0, P(z), None, P(np.ones(3)), P(n2), 2, P(np.zeros(6)), P(z), P(z), P(z), None, 8, 1)
    assert code == HELM3D_ERR_NSIZE
    n3 = np.array([2.0, 2.0, 2.0])
    code = L.helm3d_mvp(0, P(z), None, P(np.ones(3)), P(n3), 3, P(np.zeros(6)), P(z), P(z), P(z), None, 7, 1)
    assert code == HELM3D_ERR_XSIZE
    assert b"x must have" in L.helm3d_last_error()


@pytest.mark.parametrize("n", [(0.0, 2.0, 2.0), (2.5, 2.0, 2.0), (-1.0, 2.0, 2.0), (float("nan"), 2.0, 2.0)])
def test_nsize_rejects_non_positive_and_fractional_sizes(n):
    """n must hold three positive integers (Helm3dmvp_mex.c:75-79 casts them);
    the check comes before any device call."""
    import numpy as np
    import helm3d
    from helm3d._lib import HELM3D_ERR_NSIZE
    L = helm3d.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    z = np.zeros(8)
    P = lambda a: a.ctypes.data_as(dp)  # noqa: E731
    code = L.helm3d_mvp(0, P(z), None, P(np.ones(3)), P(np.array(n)), 3, P(np.zeros(6)), P(z), P(z), P(z), None,
                        8, 1)
    assert code == HELM3D_ERR_NSIZE


def test_argument_errors_without_gpu():
    """Bad variant, null operands and negative PML widths are argument errors;
    the sweep checks the rows of b (Helm3dkaczswp_mex.c Bsize)."""
    import numpy as np
    import helm3d
    from helm3d._lib import HELM3D_ERR_ARG, HELM3D_ERR_BSIZE, HELM3D_ERR_XSIZE
    L = helm3d.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    z = np.zeros(8)
    P = lambda a: a.ctypes.data_as(dp)  # noqa: E731
    n3 = np.array([2.0, 2.0, 2.0])
    h = np.ones(3)
    assert L.helm3d_mvp(9, P(z), None, P(h), P(n3), 3, P(np.zeros(6)), P(z), P(z), P(z), None, 8, 1) == HELM3D_ERR_ARG
    assert L.helm3d_mvp(0, None, None, P(h), P(n3), 3, P(np.zeros(6)), P(z), P(z), P(z), None, 8, 1) == HELM3D_ERR_ARG
    bad = np.array([1.0, -1.0, 0.0, 0.0, 0.0, 0.0])
    assert L.helm3d_mvp(0, P(z), None, P(h), P(n3), 3, P(bad), P(z), P(z), P(z), None, 8, 1) == HELM3D_ERR_ARG
    # complex-wn variant without the imaginary part
    assert L.helm3d_mvp(4, P(z), None, P(h), P(n3), 3, P(np.zeros(6)), P(z), P(z), P(z), None, 8, 1) == HELM3D_ERR_ARG
    code = L.helm3d_kaczswp(0, P(z), None, P(h), P(n3), 3, P(np.zeros(6)), 1.5, P(z), None, 8, 1, P(z), None, 7, 1,
                            P(z), P(z))
    assert code == HELM3D_ERR_BSIZE
    code = L.helm3d_kaczswp(0, P(z), None, P(h), P(n3), 3, P(np.zeros(6)), 1.5, P(z), None, 6, 1, P(z), None, 8, 1,
                            P(z), P(z))
    assert code == HELM3D_ERR_XSIZE
    # b must have as many columns as x (column s sweeps x(:,s) against b(:,s))
    code = L.helm3d_kaczswp(0, P(z), None, P(h), P(n3), 3, P(np.zeros(6)), 1.5, P(z), None, 8, 1, P(z), None, 8, 2,
                            P(z), P(z))
    assert code == HELM3D_ERR_BSIZE
    # zero columns: nothing to sweep, no device needed
    code = L.helm3d_kaczswp(0, P(z), None, P(h), P(n3), 3, P(np.zeros(6)), 1.5, None, None, 8, 0, None, None, 8, 0,
                            None, None)
    assert code == 0
    # host solves: a null context is an argument error
    assert L.helm3d_solve_host(None, 1, P(z), None, 8, 1, P(z), P(z), None, None) == HELM3D_ERR_ARG
    assert L.helm3d_set_model_host(None, P(n3), 3, P(h), P(np.zeros(6)), P(z), None) == HELM3D_ERR_ARG



def test_product_path_has_no_oracle_or_fallback():
    """The package and bench's timed legs never reach the CPU checker: no
    module of the package names it, and a missing library raises."""
    pkg = os.path.join(REPO, "slim-release-apps_amd")
    for root, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h", ".c")):
                text = open(os.path.join(root, f), errors="replace").read()
                assert not re.search(r"\boracle\b|fgmres_oracle|helm27_oracle|_ref/", text), f
    from helm3d import _lib
    saved = (_lib.LIB_PATH, _lib._lib)
    try:
        _lib.LIB_PATH = os.path.join(REPO, "no-such-dir", "libhelm3d.so")
        _lib._lib = None
        with pytest.raises(RuntimeError, match="is missing"):
            _lib.lib()
    finally:
        _lib.LIB_PATH, _lib._lib = saved
