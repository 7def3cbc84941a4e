"""numpp_amd — MI355X (gfx950) HIP device backend for NumPP's matmul and
stacked linalg.solve.

The compute lives in libnumpp_hip.so (hand-written gfx950 HIP kernels behind the
C ABI in include/numpp_hip.h); this package is the host-side mirror of NumPP's
dispatcher (backend.py, linalg.py, ufunc.py), a ctypes view of the GpuVTable provider
(vtable.py) and device-buffer plumbing (device.py). It has no CPU compute path.
"""
from . import _lib
from ._lib import BackendUnavailable, HipError, load, require_device
from .backend import Backend, CpuPathSelected, last_backend, matmul
from . import linalg, ufunc

__all__ = ["Backend", "BackendUnavailable", "CpuPathSelected", "HipError", "last_backend", "linalg", "load",
           "matmul", "require_device", "ufunc"]


def available() -> bool:
    """True when the library is built and a gfx950 device is present."""
    try:
        return bool(_lib.load().npph_available())
    except BackendUnavailable:
        return False
