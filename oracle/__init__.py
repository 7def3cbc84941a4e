"""CPU oracle for the numpp matmul / stacked-solve path (test infrastructure).

This package is the checker. It is imported only by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``; the
product library never calls it and has no CPU fallback through it.

The arithmetic lives in ``numpp_oracle.c`` (a loop-for-loop C restatement of
NumPP's portable kernels, see the citations there); this module only wraps it
with numpy arrays via ctypes. ``build()`` compiles it with the recipe in
``oracle/Makefile``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


def build() -> str:
    """Compile the C restatement (make -C oracle). Returns the .so path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.orc_lu_decompose.argtypes = [_dp, ctypes.c_int, _ip, _ip, _dp]
    lib.orc_lu_decompose.restype = ctypes.c_int
    lib.orc_lu_solve.argtypes = [_dp, ctypes.c_int, _ip, _dp, ctypes.c_int]
    lib.orc_lu_solve.restype = None
    lib.orc_solve.argtypes = [_dp, _dp, _dp, _dp, _ip, ctypes.c_int, ctypes.c_int]
    lib.orc_solve.restype = ctypes.c_int
    lib.orc_batched_solve.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _ip]
    lib.orc_batched_solve.restype = ctypes.c_int64
    lib.orc_matmul.argtypes = [ctypes.c_int64] * 3 + [_dp, _dp, _dp]
    lib.orc_matmul.restype = None
    lib.orc_matmul_f32.argtypes = [ctypes.c_int64] * 3 + [_fp, _fp, _fp]
    lib.orc_matmul_f32.restype = None
    lib.orc_dot2_entry.argtypes = [ctypes.c_int64, ctypes.c_int64, _dp, _dp, ctypes.c_int64, ctypes.c_int64]
    lib.orc_dot2_entry.restype = ctypes.c_double
    lib.orc_absdot_entry.argtypes = [ctypes.c_int64, ctypes.c_int64, _dp, _dp, ctypes.c_int64, ctypes.c_int64]
    lib.orc_absdot_entry.restype = ctypes.c_double
    lib.orc_batched_solve_numpp.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
    lib.orc_batched_solve_numpp.restype = ctypes.c_int64
    lib.orc_matmul_numpp.argtypes = [ctypes.c_int64] * 3 + [_dp, _dp, _dp]
    lib.orc_matmul_numpp.restype = None
    lib.orc_nd_copy.argtypes = [_dp, ctypes.c_int64, ctypes.c_int64]
    lib.orc_nd_copy.restype = ctypes.c_void_p
    lib.orc_reduce_f64.argtypes = [ctypes.c_int, ctypes.c_int64, _dp]
    lib.orc_reduce_f64.restype = ctypes.c_double
    lib.orc_reduce_f32.argtypes = [ctypes.c_int, ctypes.c_int64, _fp]
    lib.orc_reduce_f32.restype = ctypes.c_float
    lib.orc_ew_f64.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, _dp, _dp]
    lib.orc_ew_f64.restype = None
    lib.orc_ew_f32.argtypes = [ctypes.c_int, ctypes.c_int64, _fp, _fp, _fp]
    lib.orc_ew_f32.restype = None
    lib.orc_det.argtypes = [_dp, ctypes.c_int]
    lib.orc_det.restype = ctypes.c_double
    lib.orc_slogdet.argtypes = [_dp, ctypes.c_int, _dp, _dp]
    lib.orc_slogdet.restype = None
    lib.orc_inv.argtypes = [_dp, ctypes.c_int, _dp]
    lib.orc_inv.restype = ctypes.c_int
    lib.orc_backward_error.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
    lib.orc_backward_error.restype = None
    _lib = lib
    return lib


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


def lu_decompose(a):
    """linalg.cpp:48-71. Returns (LU, piv, steps, permsign, singular_step)."""
    a = _c(a).copy()
    n = a.shape[0]
    piv = np.zeros(n, np.int32)
    steps = np.full(n, -1, np.int32)
    sgn = ctypes.c_double(1.0)
    s = _load().orc_lu_decompose(_ptr(a), n, _ptr(piv, _ip), _ptr(steps, _ip), ctypes.byref(sgn))
    return a, piv, steps, sgn.value, s


def lu_solve(lu, piv, b):
    """linalg.cpp:73-90. b is (n,) or (n, m); returns X of the same shape."""
    lu = _c(lu)
    n = lu.shape[0]
    x = _c(b).copy()
    m = 1 if x.ndim == 1 else x.shape[1]
    piv = np.ascontiguousarray(piv, np.int32)
    _load().orc_lu_solve(_ptr(lu), n, _ptr(piv, _ip), _ptr(x), m)
    return x


class SingularMatrix(Exception):
    """Mirror of numpp::linalg_error("Singular matrix") (linalg.cpp:100)."""


def solve(a, b, with_factors=False):
    """linalg.cpp:92-103 for one (n, n) system; b is (n,) or (n, m)."""
    a = _c(a)
    b = _c(b)
    n = a.shape[0]
    m = 1 if b.ndim == 1 else b.shape[1]
    x = np.empty_like(b)
    lu = np.empty_like(a)
    steps = np.full(n, -1, np.int32)
    s = _load().orc_solve(_ptr(a), _ptr(b), _ptr(x), _ptr(lu), _ptr(steps, _ip), n, m)
    if s >= 0:
        raise SingularMatrix(f"Singular matrix (exact zero pivot at step {s})")
    if with_factors:
        return x, lu, steps
    return x


def batched_solve(a, b):
    """batched.cpp:90-100 on a (count, n, n) stack and (count, n) or (count, n, m)
    right-hand sides. Returns (X, info) with info[i] = -1 ok or the singular step."""
    a = _c(a)
    b = _c(b)
    count, n = a.shape[0], a.shape[1]
    bvec = b.ndim == a.ndim - 1
    m = 1 if bvec else b.shape[-1]
    x = np.empty_like(b)
    info = np.zeros(count, np.int32)
    _load().orc_batched_solve(count, n, m, _ptr(a), _ptr(b), _ptr(x), _ptr(info, _ip))
    return x, info


def batched_solve_numpp(a, b):
    """batched.cpp:90-100 with NumPP's per-matrix ndarray copies (the
    "NumPP-equivalent" CPU time). Returns X (same values as batched_solve)."""
    a = _c(a)
    b = _c(b)
    count, n = a.shape[0], a.shape[1]
    m = 1 if b.ndim == a.ndim - 1 else b.shape[-1]
    x = np.empty_like(b)
    _load().orc_batched_solve_numpp(count, n, m, _ptr(a), _ptr(b), _ptr(x))
    return x


def matmul_numpp(a, b):
    """backend.cpp:55-84 with the astype copies of both operands (float64)."""
    a, b = _c(a), _c(b)
    c = np.empty((a.shape[0], b.shape[1]))
    _load().orc_matmul_numpp(a.shape[0], b.shape[1], a.shape[1], _ptr(a), _ptr(b), _ptr(c))
    return c


EW_OPS = {"add": 0, "subtract": 1, "multiply": 2, "divide": 3, "negative": 10, "sqrt": 11}


def elementwise(name, a, b=None):
    """NumPP's CPU element-wise arithmetic in the array's dtype (float32/64)."""
    a = np.ascontiguousarray(a)
    f32 = a.dtype == np.float32
    a = _c(a, np.float32 if f32 else np.float64)
    b = a if b is None else _c(b, a.dtype)
    o = np.empty_like(a)
    if f32:
        _load().orc_ew_f32(EW_OPS[name], a.size, _ptr(a, _fp), _ptr(b, _fp), _ptr(o, _fp))
    else:
        _load().orc_ew_f64(EW_OPS[name], a.size, _ptr(a), _ptr(b), _ptr(o))
    return o


def reduce(name, a):
    """ufunc.cpp:450-472: sequential sum / prod in the array's dtype."""
    op = {"sum": 0, "prod": 1}[name]
    a = np.ascontiguousarray(a).ravel()
    if a.dtype == np.float32:
        return np.float32(_load().orc_reduce_f32(op, a.size, _ptr(a, _fp)))
    a = _c(a)
    return _load().orc_reduce_f64(op, a.size, _ptr(a))


def det(a):
    """linalg.cpp:393-407 for one (n, n) matrix (0.0 when singular)."""
    a = _c(a)
    return _load().orc_det(_ptr(a), a.shape[0])


def slogdet(a):
    """linalg.cpp:408-439 (real): (sign, logabsdet); (0.0, -inf) when singular."""
    a = _c(a)
    s, la = ctypes.c_double(), ctypes.c_double()
    _load().orc_slogdet(_ptr(a), a.shape[0], ctypes.byref(s), ctypes.byref(la))
    return s.value, la.value


def inv(a):
    """linalg.cpp:105-115 for one (n, n) matrix; raises SingularMatrix."""
    a = _c(a)
    n = a.shape[0]
    out = np.empty((n, n))
    s = _load().orc_inv(_ptr(a), n, _ptr(out))
    if s >= 0:
        raise SingularMatrix(f"Singular matrix (exact zero pivot at step {s})")
    return out


def matmul(a, b):
    """backend.cpp:55-84 (i-p-j loop) for float64 or float32 operands."""
    if a.dtype == np.float32 and b.dtype == np.float32:
        a, b = _c(a, np.float32), _c(b, np.float32)
        c = np.empty((a.shape[0], b.shape[1]), np.float32)
        _load().orc_matmul_f32(a.shape[0], b.shape[1], a.shape[1], _ptr(a, _fp), _ptr(b, _fp), _ptr(c, _fp))
        return c
    a, b = _c(a), _c(b)
    c = np.empty((a.shape[0], b.shape[1]), np.float64)
    _load().orc_matmul(a.shape[0], b.shape[1], a.shape[1], _ptr(a), _ptr(b), _ptr(c))
    return c


def dot2_entries(a, b, rows, cols):
    """Compensated (Dot2) entries C[rows[t], cols[t]] and (|A||B|) there."""
    a, b = _c(a), _c(b)
    k, n = a.shape[1], b.shape[1]
    lib = _load()
    ref = np.array([lib.orc_dot2_entry(n, k, _ptr(a), _ptr(b), int(i), int(j)) for i, j in zip(rows, cols)])
    mag = np.array([lib.orc_absdot_entry(n, k, _ptr(a), _ptr(b), int(i), int(j)) for i, j in zip(rows, cols)])
    return ref, mag


def backward_error(a, b, x):
    """Normwise backward error per RHS column (Dot2-accumulated residual)."""
    a, b, x = _c(a), _c(b), _c(x)
    n = a.shape[0]
    b2 = b.reshape(n, -1)
    x2 = np.ascontiguousarray(x.reshape(n, -1))
    m = b2.shape[1]
    eta = np.empty(m)
    _load().orc_backward_error(n, m, _ptr(a), _ptr(np.ascontiguousarray(b2)), _ptr(x2), _ptr(eta))
    return eta
