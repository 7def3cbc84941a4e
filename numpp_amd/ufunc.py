"""Host-side mirror of NumPP's optional GPU dispatch for ufuncs and reductions
(src/umath/ufunc.cpp:519-545, 549-552, 591, 597-598, 606-617, 653-658).

Rules followed:
  - a device path is tried when the provider's vtable is present (no target
    gate: ufuncs reach the GPU whenever one is registered), the operand dtype
    is float or complex (gpu_dtype_ok, ufunc.cpp:520), binary operands have the
    same shape, and the size is at least NUMPP_GPU_MIN (default 65536,
    ufunc.cpp:521-524);
  - binary output dtype is result_type (true divide of integers -> float64,
    arith_out, ufunc.cpp:179-200); sqrt / exp promote to float (to_float,
    ufunc.cpp:33-37); sum / prod take the device only for a full reduction
    with no axis, no keepdims and no dtype (ufunc.cpp:607-617);
  - on success last_backend() is Backend.Device; when a slot declines (exp,
    complex) or the case is ineligible, the caller's cpu_kernel (NumPP's own
    kernel in an integration) runs with last_backend() = Backend.Cpu, or
    CpuPathSelected is raised when none is given.
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib, vtable
from .backend import Backend, CpuPathSelected, _set_last

GPU_ADD, GPU_SUB, GPU_MUL, GPU_DIV = _lib.EW_ADD, _lib.EW_SUB, _lib.EW_MUL, _lib.EW_DIV
GPU_NEG, GPU_SQRT, GPU_EXP = _lib.UN_NEG, _lib.UN_SQRT, _lib.UN_EXP
GPU_SUM, GPU_PROD = _lib.RED_SUM, _lib.RED_PROD

_DT = {np.dtype(np.float32): _lib.DT_F32, np.dtype(np.float64): _lib.DT_F64,
       np.dtype(np.complex64): 12, np.dtype(np.complex128): 13}


def gpu_min() -> int:
    e = os.environ.get("NUMPP_GPU_MIN")
    if e is not None:
        try:
            v = int(e)
            if v >= 0:
                return v
        except ValueError:
            pass
    return 1 << 16


def _gpu_dtype_ok(dt) -> bool:
    return np.dtype(dt) in _DT


def _vt():
    try:
        return vtable.get()
    except _lib.BackendUnavailable:
        return None


def _cpu(cpu_kernel, name):
    _set_last(Backend.Cpu)
    if cpu_kernel is None:
        raise CpuPathSelected(f"{name} dispatched to NumPP's CPU kernel")
    return cpu_kernel()


def _arith_out(op, a, b):
    p = np.result_type(a.dtype, b.dtype)
    if op == GPU_SUB and p == np.bool_:
        raise TypeError("numpy boolean subtract is not supported")
    if op == GPU_DIV and (p.kind in "iub"):
        return np.dtype(np.float64)
    return p


def _binary(a, b, op, name, cpu_kernel):
    a, b = np.asarray(a), np.asarray(b)
    out_dt = _arith_out(op, a, b)
    vt = _vt()
    if vt is not None and _gpu_dtype_ok(out_dt) and a.shape == b.shape and a.size >= gpu_min():
        ac = np.ascontiguousarray(a, dtype=out_dt)
        bc = np.ascontiguousarray(b, dtype=out_dt)
        out = np.empty(ac.shape, out_dt)
        if vt.elementwise_binary(op, _DT[out_dt], out.size, ac.ctypes.data, bc.ctypes.data, out.ctypes.data):
            _set_last(Backend.Device)
            return out
    return _cpu(cpu_kernel, name)


def _to_float(dt):
    dt = np.dtype(dt)
    if dt.kind in "cf":
        return dt
    return np.dtype(np.float64)


def _unary(a, op, out_dt, name, cpu_kernel):
    a = np.asarray(a)
    vt = _vt()
    if vt is not None and _gpu_dtype_ok(out_dt) and a.size >= gpu_min():
        ac = np.ascontiguousarray(a, dtype=out_dt)
        out = np.empty(ac.shape, out_dt)
        if vt.elementwise_unary(op, _DT[np.dtype(out_dt)], out.size, ac.ctypes.data, out.ctypes.data):
            _set_last(Backend.Device)
            return out
    return _cpu(cpu_kernel, name)


def _reduce(a, op, axis, keepdims, dtype, name, cpu_kernel):
    a = np.asarray(a)
    vt = _vt()
    if axis is None and not keepdims and dtype is None and vt is not None and _gpu_dtype_ok(a.dtype) \
            and a.size >= gpu_min():
        ac = np.ascontiguousarray(a)
        out = np.empty((), a.dtype)
        if vt.reduce(op, _DT[a.dtype], ac.size, ac.ctypes.data, out.ctypes.data):
            _set_last(Backend.Device)
            return out
    return _cpu(cpu_kernel, name)


def add(a, b, cpu_kernel=None):
    return _binary(a, b, GPU_ADD, "add", cpu_kernel)


def subtract(a, b, cpu_kernel=None):
    return _binary(a, b, GPU_SUB, "subtract", cpu_kernel)


def multiply(a, b, cpu_kernel=None):
    return _binary(a, b, GPU_MUL, "multiply", cpu_kernel)


def divide(a, b, cpu_kernel=None):
    return _binary(a, b, GPU_DIV, "divide", cpu_kernel)


def negative(a, cpu_kernel=None):
    a = np.asarray(a)
    return _unary(a, GPU_NEG, a.dtype, "negative", cpu_kernel)


def sqrt(a, cpu_kernel=None):
    a = np.asarray(a)
    return _unary(a, GPU_SQRT, _to_float(a.dtype), "sqrt", cpu_kernel)


def exp(a, cpu_kernel=None):
    a = np.asarray(a)
    return _unary(a, GPU_EXP, _to_float(a.dtype), "exp", cpu_kernel)


def sum(a, axis=None, keepdims=False, dtype=None, cpu_kernel=None):  # noqa: A001 (NumPP's name)
    return _reduce(a, GPU_SUM, axis, keepdims, dtype, "sum", cpu_kernel)


def prod(a, axis=None, keepdims=False, dtype=None, cpu_kernel=None):
    return _reduce(a, GPU_PROD, axis, keepdims, dtype, "prod", cpu_kernel)
