"""The LapackVTable provider (gesv / inv in the output dtype,
linalg.cpp:372-382 and lapack_backend.cpp:15-34) and tensorsolve /
tensorinv (tensor_solve.cpp) on the GPU.

float32: factored and solved in float32 (LAPACK's sgesv contract, not the
portable kernel's float64): pivot rows equal LAPACK sgetrf's (scipy goldens,
first-maximum rule, well-separated pivots), X within 1e-4 relative of
LAPACK's float32 solution and with float32 backward error <= 8 n u_32; the
result differs from the float64-computed one (it really ran in float32).
float64: the FP64 LU, X within 1e-12 of the oracle. Singular input: the slot
returns false. tensorsolve / tensorinv: the reference's tridiagonal literal
(test_tensor_solve.cpp:13-39) and a seeded case against numpy goldens.
"""
import numpy as np
import pytest

import oracle
from numpp_amd import Backend, vtable
from numpp_amd import _lib
from numpp_amd import linalg as nla

pytestmark = pytest.mark.gpu


def test_sgesv_float32(hip, golden):
    g = golden("sgesv_cases.npz")
    lvt = vtable.get_lapack()
    assert lvt is not None and lvt.gesv and lvt.inv and not lvt.potrf
    for n in g["sizes"]:
        a, b = g[f"A_{n}"], g[f"B_{n}"]
        A, B = a.copy(), b.copy()
        assert lvt.gesv(int(n), 3, _lib.DT_F32, A.ctypes.data, B.ctypes.data)
        x = B
        xl = g[f"X_{n}"]
        assert np.abs(x - xl).max() <= 1e-4 * np.abs(xl).max(), n
        r = b.astype(np.float64) - a.astype(np.float64) @ x.astype(np.float64)
        eta = np.abs(r).max(0) / (np.abs(a).sum(1).max() * np.abs(x).max(0) + np.abs(b).max(0))
        assert eta.max() <= 8 * n * np.finfo(np.float32).eps / 2, (n, eta.max())
        # row order of the packed LU follows the pivots: its |L| <= 1 part and U
        # reproduce P A in float32 accuracy, and the pivot rows are LAPACK's
        lu_l = g[f"LU_{n}"]
        np.testing.assert_allclose(A, lu_l, rtol=0, atol=5e-5 * max(1.0, np.abs(lu_l).max()))
        x64 = oracle.solve(a.astype(np.float64), b.astype(np.float64))
        assert not np.array_equal(x, x64.astype(np.float32)) or n <= 3


def test_lapack_gesv_float64_and_singular(hip):
    lvt = vtable.get_lapack()
    rng = np.random.default_rng(8)
    for n in (5, 70, 300):
        a = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)
        b = rng.standard_normal((n, 2))
        A, B = a.copy(), b.copy()
        assert lvt.gesv(n, 2, _lib.DT_F64, A.ctypes.data, B.ctypes.data)
        xo = oracle.solve(a, b)
        assert np.abs(B - xo).max() <= 1e-12 * np.abs(xo).max()
        ai = np.empty_like(a)
        assert lvt.inv(n, _lib.DT_F64, a.ctypes.data, ai.ctypes.data)
        assert np.abs(ai - oracle.inv(a)).max() <= 1e-12 * np.abs(ai).max()
    z = np.zeros((4, 4))
    assert not lvt.gesv(4, 1, _lib.DT_F64, z.ctypes.data, np.ones(4).ctypes.data)
    z32 = np.zeros((4, 4), np.float32)
    assert not lvt.inv(4, _lib.DT_F32, z32.ctypes.data, z32.copy().ctypes.data)


def test_solve_dispatch_order(hip, monkeypatch):
    # a forced HIP target wins (float64 compute, float32 result); with only the
    # LAPACK provider linked, float32 is solved in float32
    rng = np.random.default_rng(9)
    a = (rng.uniform(-1, 1, (40, 40)) + 40 * np.eye(40)).astype(np.float32)
    b = rng.standard_normal(40).astype(np.float32)
    x_gpu = nla.solve(a, b, Backend.Hip)
    assert x_gpu.dtype == np.float32
    x64 = oracle.solve(a.astype(np.float64), b.astype(np.float64))
    np.testing.assert_array_equal(x_gpu, x64.astype(np.float32))
    monkeypatch.setenv("NUMPP_LAPACK_BACKEND", "hip")
    x_lap = nla.solve(a, b)
    assert x_lap.dtype == np.float32 and x_lap.shape == (40,)
    assert np.abs(x_lap - x64).max() <= 1e-5 * np.abs(x64).max()
    ai = np.empty_like(a)
    assert vtable.get_lapack().inv(40, _lib.DT_F32, a.ctypes.data, ai.ctypes.data)
    np.testing.assert_allclose(ai @ a, np.eye(40), atol=1e-5)


def test_tensorsolve_tensorinv(hip, golden):
    g = golden("tensor_cases.npz")
    x = nla.tensorsolve(g["M"], g["b"], forced=Backend.Hip)
    assert x.shape == (2, 2)
    np.testing.assert_allclose(x, g["x"], rtol=1e-12)
    np.testing.assert_allclose(g["M"].reshape(4, 4) @ x.reshape(4), g["b"].reshape(4), rtol=1e-12)
    np.testing.assert_allclose(nla.tensorinv(g["M"], 2, forced=Backend.Hip), g["Minv"], rtol=1e-12, atol=1e-15)
    x2 = nla.tensorsolve(g["A2"], g["b2"], forced=Backend.Hip)
    np.testing.assert_allclose(x2, g["x2"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(nla.tensorinv(g["A2"], 2, forced=Backend.Hip), g["A2inv"], rtol=1e-10, atol=1e-12)
