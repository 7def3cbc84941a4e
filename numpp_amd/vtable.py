"""ctypes view of the numpp::GpuVTable struct the library provides.

Field order is the reference struct (gpu_vtable.hpp:17-53 in NumPP) followed by
the appended solve slots (include/numpp_hip/gpu_vtable.hpp). Reading it through
ctypes exercises exactly the pointers a NumPP build would call.
"""
from __future__ import annotations

import ctypes

from . import _lib

_u8 = ctypes.c_uint8
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int32)
_dp = ctypes.POINTER(ctypes.c_double)

EW_BINARY_FN = ctypes.CFUNCTYPE(ctypes.c_bool, ctypes.c_int, _u8, _i64, _vp, _vp, _vp)
EW_UNARY_FN = ctypes.CFUNCTYPE(ctypes.c_bool, ctypes.c_int, _u8, _i64, _vp, _vp)
REDUCE_FN = ctypes.CFUNCTYPE(ctypes.c_bool, ctypes.c_int, _u8, _i64, _vp, _vp)
GEMM_FN = ctypes.CFUNCTYPE(ctypes.c_bool, _u8, _i64, _i64, _i64, _vp, _vp, _vp)
GESV_FN = ctypes.CFUNCTYPE(ctypes.c_bool, _u8, _i64, _i64, _i64, _vp, _vp, _vp, _vp)
GESV_REFINED_FN = ctypes.CFUNCTYPE(ctypes.c_bool, ctypes.c_int, _u8, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp)
DET_FN = ctypes.CFUNCTYPE(ctypes.c_bool, _u8, _i64, _i64, _vp, _vp, _vp)
SLOGDET_FN = ctypes.CFUNCTYPE(ctypes.c_bool, _u8, _i64, _i64, _vp, _vp, _vp, _vp)
INV_FN = ctypes.CFUNCTYPE(ctypes.c_bool, _u8, _i64, _i64, _vp, _vp, _vp)


class GpuVTable(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char_p),
        ("elementwise_binary", EW_BINARY_FN),
        ("elementwise_unary", EW_UNARY_FN),
        ("reduce", REDUCE_FN),
        ("gemm", GEMM_FN),
        ("csr_spmv", _vp),
        ("pairwise_sqdist", _vp),
        ("separable_corr1d", _vp),
        ("gesv_batched", GESV_FN),
        ("gesv_refined_batched", GESV_REFINED_FN),
        ("det_batched", DET_FN),
        ("slogdet_batched", SLOGDET_FN),
        ("inv_batched", INV_FN),
    ]


def get():
    """The provider's vtable, or None when no usable device is present."""
    p = _lib.load().npph_gpu_vtable()
    if not p:
        return None
    return GpuVTable.from_address(p)


LAPACK_GESV_FN = ctypes.CFUNCTYPE(ctypes.c_bool, ctypes.c_int, ctypes.c_int, _u8, _vp, _vp)
LAPACK_INV_FN = ctypes.CFUNCTYPE(ctypes.c_bool, ctypes.c_int, _u8, _vp, _vp)


class LapackVTable(ctypes.Structure):
    """numpp::LapackVTable (lapack_vtable.hpp:14-27 in NumPP; the provider's
    layout is include/numpp_hip/lapack_vtable.hpp)."""
    _fields_ = [
        ("gesv", LAPACK_GESV_FN),
        ("inv", LAPACK_INV_FN),
        ("potrf", _vp),
        ("gesdd", _vp),
        ("geev", _vp),
        ("heevd", _vp),
    ]


def get_lapack():
    """The provider's LapackVTable, or None when no usable device is present."""
    p = _lib.load().npph_lapack_vtable()
    if not p:
        return None
    return LapackVTable.from_address(p)
