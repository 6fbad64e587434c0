"""MI355X-native 3D frequency-domain Helmholtz hot path of SLIM-release-apps.

The product is ``libhelm3d.so`` (HIP kernels for gfx950 behind the C ABI of
``include/helm3d.h``); this package is its Python/PyTorch shim and the
host-side mirror of the reference's MATLAB interface for this path.
"""
from ._lib import Helm3dError, lib, EXPORTED  # noqa: F401
from .core import Helm3d, SolveInfo, helm3d_kaczswp_host, helm3d_mvp_host, npml_flat  # noqa: F401
from .batch import BatchSolver  # noqa: F401
from .mg import MGPool, MLGMRES  # noqa: F401

__all__ = ["Helm3d", "BatchSolver", "MLGMRES", "MGPool", "SolveInfo", "Helm3dError", "helm3d_mvp_host",
           "helm3d_kaczswp_host", "npml_flat", "lib", "EXPORTED"]
