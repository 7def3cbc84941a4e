"""Host-side mirror of NumPP's matmul dispatcher with the HIP target added.

Follows src/backend/backend.cpp (NumPP): env_target 31-41, backend_compiled
43-51, matmul 121-177, last_backend 19/118-119, with `Backend.Hip` appended to
the enum (include/numpp/backend/backend.hpp:11) and "hip" to NUMPP_GPU_TARGET.
The GPU branch calls the provider's gemm slot through the vtable. NumPP's own
CPU kernel is not part of this package: when the dispatcher selects the CPU, it
calls the `cpu_kernel` the caller passes (NumPP's cpu_matmul in an
integration) and raises CpuPathSelected when none is given.
"""
from __future__ import annotations

import enum
import os
import threading

import numpy as np

from . import _lib, vtable


class Backend(enum.Enum):
    Auto = 0
    Cpu = 1
    Blas = 2
    Device = 3
    Metal = 4
    Vulkan = 5
    Cuda = 6
    OpenCL = 7
    Hip = 8


class CpuPathSelected(RuntimeError):
    """The dispatcher chose NumPP's CPU kernel and no cpu_kernel was given."""


_tls = threading.local()


def last_backend() -> Backend:
    return getattr(_tls, "last", Backend.Cpu)


def _set_last(b: Backend):
    _tls.last = b


def env_target() -> Backend:
    s = os.environ.get("NUMPP_GPU_TARGET")
    if s is None:
        return Backend.Auto
    return {"cpu": Backend.Cpu, "metal": Backend.Metal, "vulkan": Backend.Vulkan, "cuda": Backend.Cuda,
            "opencl": Backend.OpenCL, "hip": Backend.Hip}.get(s, Backend.Auto)


def is_gpu(b: Backend) -> bool:
    return b in (Backend.Metal, Backend.Vulkan, Backend.Cuda, Backend.OpenCL, Backend.Hip)


def backend_compiled(b: Backend) -> bool:
    if b != Backend.Hip:
        return False
    try:
        _lib.load()
        return True
    except _lib.BackendUnavailable:
        return False


def gpu_available(b: Backend) -> bool:
    return backend_compiled(b) and vtable.get() is not None


def backend_name(b: Backend) -> str:
    return b.name.lower()


def result_type(a, b):
    return np.result_type(a.dtype, b.dtype)


def matmul(a, b, forced: Backend = Backend.Auto, cpu_kernel=None):
    """backend.cpp:121-177 with the HIP GPU branch."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.ndim != 2 or b.ndim != 2:
        raise ValueError("matmul requires 2-D arrays")
    if a.shape[1] != b.shape[0]:
        raise ValueError("matmul: shape mismatch")
    dt = result_type(a, b)
    m, k, n = a.shape[0], a.shape[1], b.shape[1]
    target = forced if forced != Backend.Auto else env_target()
    if is_gpu(target) and not gpu_available(target):
        raise NotImplementedError(f"backend '{backend_name(target)}' is not available in this build")
    if is_gpu(target) and dt in (np.float32, np.float64):
        vt = vtable.get()
        if vt is not None and vt.gemm:
            ac = np.ascontiguousarray(a, dtype=dt)
            bc = np.ascontiguousarray(b, dtype=dt)
            c = np.empty((m, n), dtype=dt)
            dtid = _lib.DT_F64 if dt == np.float64 else _lib.DT_F32
            if vt.gemm(dtid, m, n, k, ac.ctypes.data, bc.ctypes.data, c.ctypes.data):
                _set_last(target)
                return c
    if forced == Backend.Blas:
        raise NotImplementedError("BLAS backend not available")
    _set_last(Backend.Cpu)
    if cpu_kernel is None:
        raise CpuPathSelected("matmul dispatched to NumPP's CPU kernel")
    return cpu_kernel(a.astype(dt), b.astype(dt))
