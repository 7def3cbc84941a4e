"""Host-side mirror of numpp::linalg::solve with the HIP stacked-solve hook.

Semantics follow NumPP (src/linalg/linalg.cpp, src/linalg/batched.cpp):
  - ndim > 2 routes to the stacked solve (linalg.cpp:369); the right-hand side
    is a stack of vectors when b.ndim == a.ndim - 1 (batched.cpp:92);
  - the matrix must be square (linalg.cpp:30-33) -> LinAlgError;
  - computation is float64; the result is float32 only when the promoted input
    dtype is float32 (la_kind, linalg.cpp:20-26);
  - an exact-zero pivot raises LinAlgError("Singular matrix") (linalg.cpp:100).
The GPU hook is gated like matmul's GPU branch (backend.cpp:144-152): it runs
when the target (forced argument, else NUMPP_GPU_TARGET) is Backend.Hip, calls
the provider's gesv_batched slot once for the whole stack, and records
last_backend. Otherwise the caller's `cpu_kernel` (NumPP's per-matrix loop in
an integration) runs, or CpuPathSelected is raised.

Also exposes the device-resident API (getrf/getrs/gesv and the mixed-precision
refined solve) for callers that keep matrices in HBM.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib, vtable
from .backend import Backend, CpuPathSelected, _set_last, env_target, gpu_available, is_gpu, backend_name
from .device import DeviceArray


class LinAlgError(ValueError):
    """Mirror of numpp::linalg_error (include/numpp/core/error.hpp:43-46)."""


_DTID = {np.dtype(np.float32): _lib.DT_F32, np.dtype(np.float64): _lib.DT_F64}


def lapack_vtable():
    """The LapackVTable a NumPP build would link when the HIP provider is chosen
    for lapack_vtable() (the NUMPP_LAPACK_BACKEND=hip analogue of
    NUMPP_WITH_LAPACK), else None: the portable kernel runs."""
    import os
    if os.environ.get("NUMPP_LAPACK_BACKEND") != "hip":
        return None
    try:
        return vtable.get_lapack()
    except _lib.BackendUnavailable:
        return None


def _out_dtype(a, b):
    rt = np.result_type(a.dtype, b.dtype)
    return np.float32 if rt == np.float32 else np.float64


def solve(a, b, forced: Backend = Backend.Auto, cpu_kernel=None):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.ndim < 2 or a.shape[-1] != a.shape[-2]:
        raise LinAlgError("Last 2 dimensions of the array must be square")
    if np.iscomplexobj(a) or np.iscomplexobj(b):
        raise NotImplementedError("complex solve is served by NumPP's CPU path")
    out_dt = _out_dtype(a, b)
    lead = a.shape[:-2]
    n = a.shape[-1]
    count = int(np.prod(lead, dtype=np.int64)) if lead else 1
    bvec = (b.ndim == a.ndim - 1) if a.ndim > 2 else (b.ndim == 1)
    nrhs = 1 if bvec else b.shape[-1]
    exp_b = lead + ((n,) if bvec else (n, nrhs))
    if b.shape != exp_b:
        raise ValueError(f"solve: b has shape {b.shape}, expected {exp_b}")
    target = forced if forced != Backend.Auto else env_target()
    if is_gpu(target) and not gpu_available(target):
        raise NotImplementedError(f"backend '{backend_name(target)}' is not available in this build")
    if is_gpu(target):
        vt = vtable.get()
        if vt is not None and vt.gesv_batched:
            A = np.ascontiguousarray(a, dtype=np.float64).reshape(count, n, n).copy()
            B = np.ascontiguousarray(b, dtype=np.float64).reshape(count, n, nrhs).copy()
            info = np.zeros(count, np.int32)
            if vt.gesv_batched(_lib.DT_F64, count, n, nrhs, A.ctypes.data, B.ctypes.data, None,
                               info.ctypes.data):
                _set_last(target)
                if np.any(info != 0):
                    raise LinAlgError("Singular matrix")
                return B.reshape(exp_b).astype(out_dt, copy=False)
    # A linked LAPACK provider (linalg.cpp:372-382): 2-D systems only, computed
    # in the output dtype; a false return (e.g. singular) falls through.
    lvt = lapack_vtable() if a.ndim == 2 else None
    if lvt is not None and lvt.gesv:
        A = np.ascontiguousarray(a, dtype=out_dt).copy()
        B = np.ascontiguousarray(b, dtype=out_dt).reshape(n, nrhs).copy()
        if lvt.gesv(n, nrhs, _DTID[np.dtype(out_dt)], A.ctypes.data, B.ctypes.data):
            _set_last(Backend.Cpu)
            return B.reshape(b.shape)
    _set_last(Backend.Cpu)
    if cpu_kernel is None:
        raise CpuPathSelected("solve dispatched to NumPP's CPU kernel")
    return cpu_kernel(a, b).astype(out_dt, copy=False)


def _square_stack(a):
    """Shape checks of require_square (linalg.cpp:30-33) and batched split
    (batched.cpp:22-30): (lead, count, n)."""
    if a.ndim < 2 or a.shape[-1] != a.shape[-2]:
        raise LinAlgError("Last 2 dimensions of the array must be square")
    if np.iscomplexobj(a):
        raise NotImplementedError("complex input is served by NumPP's CPU path")
    lead = a.shape[:-2]
    count = int(np.prod(lead, dtype=np.int64)) if lead else 1
    return lead, count, a.shape[-1]


def _gpu_target(forced):
    target = forced if forced != Backend.Auto else env_target()
    if is_gpu(target) and not gpu_available(target):
        raise NotImplementedError(f"backend '{backend_name(target)}' is not available in this build")
    return target


def _cpu(name, cpu_kernel, *args):
    _set_last(Backend.Cpu)
    if cpu_kernel is None:
        raise CpuPathSelected(f"{name} dispatched to NumPP's CPU kernel")
    return cpu_kernel(*args)


def det(a, forced: Backend = Backend.Auto, cpu_kernel=None):
    """numpp::linalg::det (linalg.cpp:393-407; stacks: batched::det,
    batched.cpp:76) with the HIP det_batched slot: 0 for an exact-zero pivot,
    output dtype float32 only for float32 input."""
    a = np.asarray(a)
    lead, count, n = _square_stack(a)
    out_dt = _out_dtype(a, a)
    target = _gpu_target(forced)
    if is_gpu(target):
        vt = vtable.get()
        if vt is not None and vt.det_batched:
            A = np.ascontiguousarray(a, dtype=np.float64)
            d = np.empty(count)
            if vt.det_batched(_lib.DT_F64, count, n, A.ctypes.data, d.ctypes.data, None):
                _set_last(target)
                return d.reshape(lead).astype(out_dt, copy=False) if lead else out_dt(d[0])
    return _cpu("det", cpu_kernel, a)


def slogdet(a, forced: Backend = Backend.Auto, cpu_kernel=None):
    """numpp::linalg::slogdet (linalg.cpp:408-439; stacks: batched.cpp:102-108):
    (sign, logabsdet), (0, -inf) for an exact-zero pivot."""
    a = np.asarray(a)
    lead, count, n = _square_stack(a)
    out_dt = _out_dtype(a, a)
    target = _gpu_target(forced)
    if is_gpu(target):
        vt = vtable.get()
        if vt is not None and vt.slogdet_batched:
            A = np.ascontiguousarray(a, dtype=np.float64)
            sg = np.empty(count)
            la = np.empty(count)
            if vt.slogdet_batched(_lib.DT_F64, count, n, A.ctypes.data, sg.ctypes.data, la.ctypes.data, None):
                _set_last(target)
                if lead:
                    return sg.reshape(lead).astype(out_dt, copy=False), la.reshape(lead).astype(out_dt, copy=False)
                return out_dt(sg[0]), out_dt(la[0])
    return _cpu("slogdet", cpu_kernel, a)


def inv(a, forced: Backend = Backend.Auto, cpu_kernel=None):
    """numpp::linalg::inv (linalg.cpp:105-115, 386-392; stacks: batched::inv,
    batched.cpp:75) with the HIP inv_batched slot; LinAlgError("Singular
    matrix") for an exact-zero pivot."""
    a = np.asarray(a)
    lead, count, n = _square_stack(a)
    out_dt = _out_dtype(a, a)
    target = _gpu_target(forced)
    if is_gpu(target):
        vt = vtable.get()
        if vt is not None and vt.inv_batched:
            A = np.ascontiguousarray(a, dtype=np.float64)
            out = np.empty((count, n, n))
            info = np.zeros(count, np.int32)
            if vt.inv_batched(_lib.DT_F64, count, n, A.ctypes.data, out.ctypes.data, info.ctypes.data):
                _set_last(target)
                if np.any(info != 0):
                    raise LinAlgError("Singular matrix")
                return out.reshape(a.shape).astype(out_dt, copy=False)
    return _cpu("inv", cpu_kernel, a)


def tensorsolve(a, b, axes=None, forced: Backend = Backend.Auto, cpu_kernel=None):
    """numpp::linalg::tensorsolve (tensor_solve.cpp:10-32): move the named axes
    of a to the end, square the trailing axes against b's shape, and solve the
    (prod, prod) system through solve() (which takes the GPU hook)."""
    a = np.asarray(a)
    b = np.asarray(b)
    if axes:
        allaxes = list(range(a.ndim))
        for k in axes:
            allaxes.remove(k)
            allaxes.append(k)
        a = a.transpose(allaxes)
    if a.ndim < b.ndim:
        raise LinAlgError("tensorsolve: a.ndim must be >= b.ndim")
    oldshape = a.shape[b.ndim:] if a.ndim > b.ndim else ()
    prod = int(np.prod(oldshape, dtype=np.int64))
    if a.size != prod * prod:
        raise LinAlgError("tensorsolve: prod(a.shape[b.ndim:]) must equal prod(a.shape[:b.ndim])")
    A = np.ascontiguousarray(a).reshape(prod, prod)
    B = np.ascontiguousarray(b).reshape(prod)
    kern = None if cpu_kernel is None else (lambda x, y: cpu_kernel(x, y))
    return solve(A, B, forced, kern).reshape(oldshape)


def tensorinv(a, ind=2, forced: Backend = Backend.Auto, cpu_kernel=None):
    """numpp::linalg::tensorinv (tensor_solve.cpp:34-49): inverse of a reshaped
    to (prod(shape[ind:]), prod(shape[:ind])) through inv(), reshaped to
    shape[ind:] + shape[:ind]."""
    a = np.asarray(a)
    if ind <= 0 or ind > a.ndim:
        raise ValueError("Invalid ind argument")
    sh = a.shape
    prod = int(np.prod(sh[ind:], dtype=np.int64))
    lead = int(np.prod(sh[:ind], dtype=np.int64))
    A = np.ascontiguousarray(a).reshape(prod, lead)
    return inv(A, forced, cpu_kernel).reshape(sh[ind:] + sh[:ind])


# ---------------------------------------------------------------------------
# Device-resident API


def _ptr(x):
    return None if x is None else x.ptr


def gesv_batched_dev(A: DeviceArray, B: DeviceArray, ipiv: DeviceArray | None = None,
                     info: DeviceArray | None = None, stream=None):
    """In place: A (count, n, n) -> LU, B (count, n, nrhs) -> X (FP64)."""
    lib = _lib.require_device()
    count, n, _ = A.shape
    nrhs = B.shape[2] if len(B.shape) == 3 else 1
    _lib.check(lib.npph_gesv_batched_dev(count, n, nrhs, A.ptr, B.ptr, _ptr(ipiv), _ptr(info), stream))


def getrf_batched_dev(A: DeviceArray, ipiv: DeviceArray, info: DeviceArray, stream=None):
    lib = _lib.require_device()
    count, n, _ = A.shape
    _lib.check(lib.npph_getrf_batched_dev(count, n, A.ptr, n * n, ipiv.ptr, info.ptr, stream))


def getrs_batched_dev(LU: DeviceArray, ipiv: DeviceArray, B: DeviceArray, stream=None):
    lib = _lib.require_device()
    count, n, _ = LU.shape
    nrhs = B.shape[2] if len(B.shape) == 3 else 1
    _lib.check(lib.npph_getrs_batched_dev(count, n, nrhs, LU.ptr, n * n, ipiv.ptr, B.ptr, n * nrhs, stream))


PRECISIONS = {"fp64": _lib.PREC_FP64, "fp32": _lib.PREC_FP32, "bf16": _lib.PREC_BF16}


def refine_opts(precision="fp32", tol=1e-12, max_iter=10, fallback=True, bf16_terms=6):
    return _lib.RefineOpts(PRECISIONS[precision] if isinstance(precision, str) else int(precision),
                           int(max_iter), float(tol), 1 if fallback else 0, int(bf16_terms))


def det_batched_dev(A: DeviceArray, det: DeviceArray | None = None, sign: DeviceArray | None = None,
                    logabsdet: DeviceArray | None = None, info: DeviceArray | None = None, stream=None):
    """In place: A (count, n, n) -> LU; per-matrix det / sign / logabsdet (FP64)."""
    lib = _lib.require_device()
    count, n, _ = A.shape
    _lib.check(lib.npph_det_batched_dev(count, n, A.ptr, _ptr(det), _ptr(sign), _ptr(logabsdet), _ptr(info),
                                        stream))


def inv_batched_dev(A: DeviceArray, Ainv: DeviceArray, info: DeviceArray | None = None, stream=None):
    """A (count, n, n) -> LU in place; Ainv receives the inverses (FP64)."""
    lib = _lib.require_device()
    count, n, _ = A.shape
    _lib.check(lib.npph_inv_batched_dev(count, n, A.ptr, Ainv.ptr, _ptr(info), stream))


def gesv_refined_dev(A: DeviceArray, B: DeviceArray, X: DeviceArray, berr: DeviceArray | None = None,
                     iters: DeviceArray | None = None, info: DeviceArray | None = None, opts=None,
                     stream=None):
    """Mixed-precision factorisation + FP64 refinement, device-resident."""
    lib = _lib.require_device()
    count, n, _ = A.shape
    nrhs = B.shape[2] if len(B.shape) == 3 else 1
    o = opts or refine_opts()
    _lib.check(lib.npph_gesv_refined_batched_dev(ctypes.byref(o), count, n, nrhs, A.ptr, B.ptr, X.ptr,
                                                 _ptr(berr), _ptr(iters), _ptr(info), stream))


def solve_refined(a, b, precision="fp32", tol=1e-12, max_iter=10, fallback=True, bf16_terms=6):
    """Host convenience: returns (x, berr, iters, info) for a (count, n, n) stack."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    squeeze = b.ndim == a.ndim - 1
    if squeeze:
        b = b[..., None]
    count, n, _ = a.shape
    nrhs = b.shape[2]
    dA, dB = DeviceArray.from_host(a), DeviceArray.from_host(b)
    dX = DeviceArray((count, n, nrhs))
    dE = DeviceArray((count, nrhs))
    dI = DeviceArray((count,), np.int32)
    dF = DeviceArray((count,), np.int32)
    gesv_refined_dev(dA, dB, dX, dE, dI, dF, refine_opts(precision, tol, max_iter, fallback, bf16_terms))
    x = dX.to_host()
    return (x[..., 0] if squeeze else x), dE.to_host(), dI.to_host(), dF.to_host()
