"""ctypes binding of libnumpp_hip.so (the C ABI declared in include/numpp_hip.h).

Loading fails loudly: there is no CPU fallback in this package. A missing
library raises ``BackendUnavailable`` with the build command to run.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnumpp_hip.so")

NPPH_OK = 0
NPPH_ERR_NO_DEVICE = -1
NPPH_ERR_INVALID = -2
NPPH_ERR_HIP = -3
NPPH_ERR_OOM = -4
NPPH_ERR_UNSUPPORTED = -5
NPPH_ERR_BUDGET = -6
NPPH_ERR_NOT_FOUND = -7
NPPH_ERR_SINGULAR = -8

DT_F32 = 10
DT_F64 = 11

EW_ADD, EW_SUB, EW_MUL, EW_DIV = 0, 1, 2, 3
UN_NEG, UN_SQRT, UN_EXP = 0, 1, 2
RED_SUM, RED_PROD = 0, 1

PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16 = 2


class BackendUnavailable(RuntimeError):
    """The HIP library is not built or no gfx950 device is present."""


class HipError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"npph status {status}: {msg}")
        self.status = status


class ServeConfig(ctypes.Structure):
    _fields_ = [
        ("budget_bytes", ctypes.c_double),
        ("workers", ctypes.c_int),
        ("max_batch", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("max_iter", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("fallback_fp64", ctypes.c_int),
        ("bf16_terms", ctypes.c_int),
    ]


class RefineOpts(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int),
        ("max_iter", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("fallback_fp64", ctypes.c_int),
        ("bf16_terms", ctypes.c_int),
    ]


_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_int = ctypes.c_int
_dbl = ctypes.c_double
_flt = ctypes.c_float

# name -> (restype, argtypes). Every symbol declared in include/numpp_hip.h.
SIGNATURES = {
    "npph_init": (_int, []),
    "npph_available": (_int, []),
    "npph_last_error": (ctypes.c_char_p, []),
    "npph_device_name": (ctypes.c_char_p, []),
    "npph_cu_count": (_int, []),
    "npph_device_id": (_int, []),
    "npph_malloc": (_int, [ctypes.POINTER(_vp), ctypes.c_uint64]),
    "npph_free": (_int, [_vp]),
    "npph_memcpy_htod": (_int, [_vp, _vp, ctypes.c_uint64, _vp]),
    "npph_memcpy_dtoh": (_int, [_vp, _vp, ctypes.c_uint64, _vp]),
    "npph_memcpy_dtod": (_int, [_vp, _vp, ctypes.c_uint64, _vp]),
    "npph_memset": (_int, [_vp, _int, ctypes.c_uint64, _vp]),
    "npph_stream_create": (_int, [ctypes.POINTER(_vp)]),
    "npph_stream_destroy": (_int, [_vp]),
    "npph_stream_sync": (_int, [_vp]),
    "npph_device_sync": (_int, []),
    "npph_event_create": (_int, [ctypes.POINTER(_vp)]),
    "npph_event_destroy": (_int, [_vp]),
    "npph_event_record": (_int, [_vp, _vp]),
    "npph_event_elapsed_ms": (_int, [_vp, _vp, ctypes.POINTER(_flt)]),
    "npph_elementwise_binary": (_int, [_int, _int, _i64, _vp, _vp, _vp]),
    "npph_elementwise_unary": (_int, [_int, _int, _i64, _vp, _vp]),
    "npph_reduce": (_int, [_int, _int, _i64, _vp, _vp]),
    "npph_elementwise_binary_dev": (_int, [_int, _int, _i64, _vp, _vp, _vp, _vp]),
    "npph_elementwise_unary_dev": (_int, [_int, _int, _i64, _vp, _vp, _vp]),
    "npph_reduce_dev": (_int, [_int, _int, _i64, _vp, _vp, _vp]),
    "npph_gemm": (_int, [_int, _i64, _i64, _i64, _vp, _vp, _vp]),
    "npph_dgemm_dev": (_int, [_i64, _i64, _i64, _dbl, _vp, _i64, _i64, _vp, _i64, _i64, _dbl, _vp, _i64, _i64, _i64, _vp]),
    "npph_sgemm_dev": (_int, [_i64, _i64, _i64, _flt, _vp, _i64, _i64, _vp, _i64, _i64, _flt, _vp, _i64, _i64, _i64, _vp]),
    "npph_bf16split_update_dev": (_int, [_i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _int, _vp]),
    "npph_gesv_batched": (_int, [_int, _i64, _i64, _i64, _vp, _vp, _vp, _vp]),
    "npph_getrf_batched_dev": (_int, [_i64, _i64, _vp, _i64, _vp, _vp, _vp]),
    "npph_getrs_batched_dev": (_int, [_i64, _i64, _i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    "npph_gesv_batched_dev": (_int, [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp]),
    "npph_gesv_batched_dev_noalloc": (_int, [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp]),
    "npph_det_batched": (_int, [_int, _i64, _i64, _vp, _vp, _vp, _vp, _vp]),
    "npph_inv_batched": (_int, [_int, _i64, _i64, _vp, _vp, _vp]),
    "npph_det_batched_dev": (_int, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "npph_inv_batched_dev": (_int, [_i64, _i64, _vp, _vp, _vp, _vp]),
    "npph_gesv_refined_batched_dev": (_int, [ctypes.POINTER(RefineOpts), _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "npph_vbatch_plan_create": (_int, [_i64, _vp, _vp, ctypes.POINTER(RefineOpts), _int, ctypes.POINTER(_vp)]),
    "npph_vbatch_plan_destroy": (_int, [_vp]),
    "npph_vbatch_plan_info": (_int, [_vp, ctypes.POINTER(_int), _vp, _vp, _vp, _int]),
    "npph_vbatch_execute": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _int, _vp]),
    "npph_server_create": (_int, [ctypes.POINTER(ServeConfig), ctypes.POINTER(_vp)]),
    "npph_server_destroy": (_int, [_vp]),
    "npph_server_submit": (_int, [_vp, _i64, _i64, _vp, _vp, _vp, _vp, ctypes.c_uint64, _int, ctypes.POINTER(_i64)]),
    "npph_server_wait": (_int, [_vp, _i64, _vp, _vp, ctypes.POINTER(_dbl), ctypes.POINTER(_dbl)]),
    "npph_server_stats": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "npph_server_replay": (_int, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_dbl)]),
    "npph_prof_enable": (_int, [_int]),
    "npph_prof_reset": (_int, []),
    "npph_prof_get": (_int, [_int, ctypes.POINTER(_dbl), ctypes.POINTER(_dbl), ctypes.POINTER(_dbl), ctypes.POINTER(_i64)]),
    "npph_gpu_vtable": (_vp, []),
    "npph_lapack_vtable": (_vp, []),
    "npph_lapack_gesv": (_int, [_int, _int, _int, _vp, _vp]),
    "npph_lapack_inv": (_int, [_int, _int, _vp, _vp]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load libnumpp_hip.so and declare every C-ABI signature (no device use)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BackendUnavailable(f"{LIB_PATH} is not built; run `python -m numpp_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().npph_last_error()
    return msg.decode() if msg else ""


def check(status: int) -> None:
    if status != NPPH_OK:
        raise HipError(status, last_error())


def require_device() -> ctypes.CDLL:
    """The library with a usable gfx950 device, else BackendUnavailable."""
    lib = load()
    if not lib.npph_available():
        raise BackendUnavailable(f"no usable gfx950 device: {last_error()}")
    return lib
