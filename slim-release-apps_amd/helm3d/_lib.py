"""ctypes binding of libhelm3d.so (the C ABI declared in include/helm3d.h).

torch is imported first so that the HIP runtime torch ships (SONAME
libamdhip64.so.7) is the one the library binds to: device pointers and
streams are then shared between torch tensors and the library.  The library
is required: every entry point raises if it is missing, there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  -- must precede loading the HIP library

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhelm3d.so")

HELM3D_OK = 0
HELM3D_ERR_NSIZE = 1
HELM3D_ERR_XSIZE = 2
HELM3D_ERR_ARG = 3
HELM3D_ERR_HIP = 4
HELM3D_ERR_NOMODEL = 5
HELM3D_ERR_NOMEM = 6
HELM3D_ERR_NOTCONV = 7
HELM3D_ERR_BSIZE = 8

HELM3D_FORW = 0
HELM3D_ADJ = 1
HELM3D_DERIV = 2
HELM3D_WNC = 4

HELM3D_C128 = 0
HELM3D_C64 = 1

HELM3D_OPT_GRAPHS = 1

HELM3D_BATCH_PLAIN = 0
HELM3D_BATCH_MIXED = 1
HELM3D_BATCH_MAX_ITEMS = 1024  # items per helm3d_batch_solve call

HELM3D_NOALLOC_MAX_AXIS = 4094  # points per axis of helm3d_apply_noalloc

HELM3D_GRID_PHYS = 0
HELM3D_GRID_PML = 1

HELM3D_UNITS = {"m/s": 0, "s2/m2": 1, "s2/km2": 2}

# MEX error identifiers of Helm3dmvp_mex.c:77,83 (the Xsize id names the sweep
# gateway in the reference too).
MEX_ERR_IDS = {
    HELM3D_ERR_NSIZE: "SLIM_release_apps:tools:algorithms:ThreeDFreqModeling:Helm3dmvp_mex:Nsize",
    HELM3D_ERR_XSIZE: "SLIM_release_apps:tools:algorithms:ThreeDFreqModeling:Helm3dkaczswp_mex:Xsize",
}
# identifiers of the sweep gateway (Helm3dkaczswp_mex.c:150,159,164)
MEX_SWEEP_ERR_IDS = {
    HELM3D_ERR_NSIZE: "SLIM_release_apps:tools:algorithms:ThreeDFreqModeling:Helm3dkaczswp_mex:Nsize",
    HELM3D_ERR_XSIZE: "SLIM_release_apps:tools:algorithms:ThreeDFreqModeling:Helm3dkaczswp_mex:Xsize",
    HELM3D_ERR_BSIZE: "SLIM_release_apps:tools:algorithms:ThreeDFreqModeling:Helm3dkaczswp_mex:Bsize",
}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


class SolveOpts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("maxinnerit", ctypes.c_int), ("maxit", ctypes.c_longlong),
                ("use_x0", ctypes.c_int), ("inner_tol", ctypes.c_double), ("inner_m", ctypes.c_int),
                ("inner_maxit", ctypes.c_int), ("inner_solver", ctypes.c_int)]


HELM3D_INNER_BICGSTAB = 0
HELM3D_INNER_GMRES = 1


class SolveInfo(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_longlong), ("inner_iters", ctypes.c_longlong), ("cycles", ctypes.c_longlong),
                ("relres", ctypes.c_double), ("quasi_relres", ctypes.c_double), ("converged", ctypes.c_int)]


class Helm3dError(RuntimeError):
    def __init__(self, code: int, msg: str, ids=None):
        super().__init__(f"helm3d error {code}: {msg}")
        self.code = code
        self.identifier = (ids if ids is not None else MEX_ERR_IDS).get(code, "")


_lib = None

# (name, restype, argtypes) of every exported entry point
_SIGS = [
    ("helm3d_version", ctypes.c_char_p, []),
    ("helm3d_last_error", ctypes.c_char_p, []),
    ("helm3d_mvp", ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_longlong, _dp, _dp, _dp, _dp, _dp,
                                  ctypes.c_longlong, ctypes.c_int]),
    ("helm3d_create", ctypes.c_int, [ctypes.c_int, _vp, ctypes.POINTER(_vp)]),
    ("helm3d_destroy", ctypes.c_int, [_vp]),
    ("helm3d_set_stream", ctypes.c_int, [_vp, _vp]),
    ("helm3d_set_model", ctypes.c_int, [_vp, _ip, _dp, _ip, _vp, ctypes.c_int]),
    ("helm3d_numel", ctypes.c_longlong, [_vp]),
    ("helm3d_apply", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    ("helm3d_apply_noalloc", ctypes.c_int, [ctypes.c_int, _ip, _dp, _ip, _vp, ctypes.c_int, _vp, _vp, _vp]),
    ("helm3d_apply_deriv", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]),
    ("helm3d_solve_opts_default", None, [ctypes.POINTER(SolveOpts)]),
    ("helm3d_solve", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.POINTER(SolveOpts),
                                    ctypes.POINTER(SolveInfo), _dp, ctypes.c_longlong,
                                    ctypes.POINTER(ctypes.c_longlong)]),
    ("helm3d_solve_mixed", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.POINTER(SolveOpts),
                                          ctypes.POINTER(SolveInfo)]),
    ("helm3d_solve_bicgstab", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.POINTER(SolveOpts),
                                             ctypes.POINTER(SolveInfo)]),
    ("helm3d_fwi_grad_accum", ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int, _vp]),
    ("helm3d_param2wavenum", ctypes.c_int, [_vp, _ip, _ip, _vp, ctypes.c_int, ctypes.c_double, _vp, _vp, _vp]),
    ("helm3d_extension_adjoint", ctypes.c_int, [_vp, _ip, _ip, _vp, _vp, ctypes.c_int]),
    ("helm3d_csr_apply", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_double, _vp, _vp,
                                        ctypes.c_int]),
    ("helm3d_batch_create", ctypes.c_int, [ctypes.c_int, _vp, ctypes.POINTER(_vp)]),
    ("helm3d_batch_destroy", ctypes.c_int, [_vp]),
    ("helm3d_batch_set_stream", ctypes.c_int, [_vp, _vp]),
    ("helm3d_batch_solve", ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.c_int,
                                          ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int,
                                          ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]),
    ("helm3d_batch_item_bytes", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]),
    ("helm3d_kacz_sweep", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _vp, _vp]),
    ("helm3d_kaczswp", ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_longlong, _dp, ctypes.c_double,
                                      _dp, _dp, ctypes.c_longlong, ctypes.c_longlong, _dp, _dp, ctypes.c_longlong,
                                      ctypes.c_longlong, _dp, _dp]),
    ("helm3d_set_model_host", ctypes.c_int, [_vp, _dp, ctypes.c_longlong, _dp, _dp, _dp, _dp]),
    ("helm3d_solve_host", ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, ctypes.c_longlong, ctypes.c_longlong, _dp,
                                         _dp, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]),
    ("helm3d_solve_cgmn", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.POINTER(SolveOpts),
                                         ctypes.POINTER(SolveInfo)]),
    ("helm3d_solve_crmn", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.POINTER(SolveOpts),
                                         ctypes.POINTER(SolveInfo)]),
    ("helm3d_mg_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                        ctypes.POINTER(_vp)]),
    ("helm3d_mg_destroy", ctypes.c_int, [_vp]),
    ("helm3d_mg_set_stream", ctypes.c_int, [_vp, _vp]),
    ("helm3d_mg_transfer", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    ("helm3d_solve_mg", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.POINTER(SolveOpts), _ip,
                                       ctypes.POINTER(SolveInfo)]),
    ("helm3d_solve_mg_mixed", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.POINTER(SolveOpts), _ip,
                                             ctypes.POINTER(SolveInfo)]),
    ("helm3d_set_launch", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("helm3d_set_option", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_longlong]),
    ("helm3d_get_launch", ctypes.c_int, [_vp, _ip, _ip, _ip, _ip, _ip]),
    ("helm3d_autotune", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _ip, _dp]),
    ("helm3d_profile", ctypes.c_int, [_vp, ctypes.c_int]),
    ("helm3d_profile_read", ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_longlong)]),
]
PROF_KINDS = {"stencil": 0, "dots": 1, "update": 2, "combine": 3, "reduce": 4}
for _v in ("forw", "adj", "forw_deriv", "adj_deriv", "forw_wnc", "adj_wnc"):
    _SIGS.append((f"helm3d_do_Hmvp_{_v}", ctypes.c_int, [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                                        ctypes.c_int]))

EXPORTED = tuple(name for name, _, _ in _SIGS)


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "or `make -C slim-release-apps_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code: int, allow=(), ids=None):
    if code != HELM3D_OK and code not in allow:
        raise Helm3dError(code, lib().helm3d_last_error().decode(), ids)
    return code
