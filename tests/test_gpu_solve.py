"""GPU stacked-solve parity against the oracle's restatement of NumPP's LU
(linalg.cpp:48-90) and batched loop (batched.cpp:90-100).

Tolerance (stated): ||x_gpu - x_cpu||_inf <= 1e-12 * kappa_est * ||x_cpu||_inf per
system, with kappa_est = 1 for the diagonally dominant workload and the
measured 1-norm condition estimate otherwise; pivots bit-exact where the
oracle certifies every pivot choice by a relative gap of at least 1e-6; the
singular rule (exact zero) gives the same first step as the oracle.
"""
import numpy as np
import pytest

import oracle
from conftest import config1_inputs
from numpp_amd import Backend, last_backend
from numpp_amd import _lib
from numpp_amd import linalg as nla
from numpp_amd.device import DeviceArray

pytestmark = pytest.mark.gpu


def gesv_host(a, b):
    """The vtable slot path: (LU, X, ipiv, info) from npph_gesv_batched."""
    a = np.ascontiguousarray(a, np.float64).copy()
    b = np.ascontiguousarray(b, np.float64).copy()
    count, n = a.shape[0], a.shape[1]
    nrhs = b.shape[2]
    ipiv = np.zeros((count, n), np.int32)
    info = np.zeros(count, np.int32)
    _lib.check(_lib.load().npph_gesv_batched(_lib.DT_F64, count, n, nrhs, a.ctypes.data, b.ctypes.data,
                                             ipiv.ctypes.data, info.ctypes.data))
    return a, b, ipiv, info


def pivot_gaps(a):
    """Smallest relative gap between the chosen pivot and the runner-up over all
    steps of the oracle's elimination (certifies pivot choices)."""
    w = np.array(a, dtype=np.float64, copy=True)
    n = w.shape[0]
    gap = np.inf
    for k in range(n - 1):
        col = np.abs(w[k:, k])
        i = int(np.argmax(col))
        best = col[i]
        rest = np.delete(col, i)
        if rest.size and best > 0:
            gap = min(gap, (best - rest.max()) / best)
        p = k + i
        w[[k, p]] = w[[p, k]]
        f = w[k + 1:, k] / w[k, k]
        w[k + 1:, k + 1:] -= np.outer(f, w[k, k + 1:])
        w[k + 1:, k] = f
    return gap


def test_config1_stack(hip, golden):
    """256 x 64 x 64 (A = U[-1,1) + 64 I, b ~ N(0,1), seed 0): the one-wave kernel."""
    a, b = config1_inputs()
    lu, x, ipiv, info = gesv_host(a, b[..., None])
    assert np.all(info == 0)
    xo, oinfo = oracle.batched_solve(a, b)
    x = x[..., 0]
    assert np.abs(x - xo).max() <= 1e-12 * np.abs(xo).max()
    # diagonally dominant: no row swaps at all (every step keeps its row)
    assert np.array_equal(ipiv, np.tile(np.arange(64, dtype=np.int32), (256, 1)))
    xg = golden("config1_x.npy")
    assert np.abs(x - xg).max() <= 1e-12 * np.abs(xg).max()


def test_golden_cases_pivots_and_factors(hip, golden):
    g = golden("gesv_cases.npz")
    for n in g["sizes"]:
        a, bb = g[f"A_{n}"], g[f"B_{n}"]
        lu, x, ipiv, info = gesv_host(a[None], bb[None])
        assert info[0] == 0
        xo, luo, steps = oracle.solve(a, bb, with_factors=True)
        kappa = np.linalg.cond(a, 1)
        assert np.abs(x[0] - xo).max() <= 1e-13 * kappa * np.abs(xo).max(), n
        if pivot_gaps(a) > 1e-6:
            np.testing.assert_array_equal(ipiv[0], steps)
            np.testing.assert_array_equal(ipiv[0], g[f"ipiv_{n}"])
            assert np.abs(lu[0] - luo).max() <= 1e-12 * kappa * np.abs(luo).max()


@pytest.mark.parametrize("n", [1, 2, 7, 8, 9, 16, 31, 33, 63, 64, 65, 100, 127, 128, 129, 200, 256, 300, 511, 513])
@pytest.mark.parametrize("nrhs", [1, 3, 9])
def test_sizes_vs_oracle(hip, n, nrhs):
    rng = np.random.default_rng(n * 31 + nrhs)
    count = 3
    a = rng.uniform(-1, 1, (count, n, n))
    a[1] += np.diag(rng.uniform(1, 2, n))  # one dominant, others general
    b = rng.standard_normal((count, n, nrhs))
    lu, x, ipiv, info = gesv_host(a, b)
    assert np.all(info == 0)
    for i in range(count):
        xo, luo, steps = oracle.solve(a[i], b[i], with_factors=True)
        kappa = np.linalg.cond(a[i], 1)
        assert np.abs(x[i] - xo).max() <= 1e-13 * kappa * np.abs(xo).max(), (n, nrhs, i)
        if pivot_gaps(a[i]) > 1e-6:
            np.testing.assert_array_equal(ipiv[i], steps)


def test_large_blocked_2048(hip):
    rng = np.random.default_rng(21)
    n = 2048
    a = rng.uniform(-1, 1, (1, n, n))
    b = rng.standard_normal((1, n, 2))
    lu, x, ipiv, info = gesv_host(a, b)
    assert info[0] == 0
    xo = oracle.solve(a[0], b[0])
    eta = oracle.backward_error(a[0], b[0], x[0])
    assert np.all(eta < 1e-14)
    assert np.abs(x[0] - xo).max() <= 1e-13 * np.linalg.cond(a[0], 1) * np.abs(xo).max()


def test_grouped_stack_1024(hip):
    # 16 FP64 systems of order 1024: the factorisation runs as four concurrent
    # groups of four matrices on their own streams. Pivots and factors of one
    # matrix per group equal the oracle's; every system's backward error is
    # FP64's, and the FP32-refined solve of the same stack reaches 1e-12.
    rng = np.random.default_rng(23)
    count, n = 16, 1024
    a = rng.uniform(-1, 1, (count, n, n))
    b = rng.standard_normal((count, n, 3))
    lu, x, ipiv, info = gesv_host(a, b)
    assert np.all(info == 0)
    for i in range(count):
        assert np.all(oracle.backward_error(a[i], b[i], x[i]) < 1e-14), i
    for i in (1, 6, 11, 12):
        xo, luo, steps = oracle.solve(a[i], b[i], with_factors=True)
        if pivot_gaps(a[i]) > 1e-6:
            np.testing.assert_array_equal(ipiv[i], steps)
            assert np.abs(lu[i] - luo).max() <= 1e-10 * np.abs(luo).max(), i
        assert np.abs(x[i] - xo).max() <= 1e-13 * np.linalg.cond(a[i], 1) * np.abs(xo).max(), i
    xr, berr, iters, rinfo = nla.solve_refined(a, b, precision="fp32")
    assert np.all(rinfo == 0) and np.all(berr <= 1e-12)
    for i in (0, 5, 10, 15):
        assert np.all(oracle.backward_error(a[i], b[i], xr[i]) <= 1e-12), i


def test_large_blocked_pivots_2100(hip):
    # n = 2100: the first panels span more than 2048 rows and the last one is
    # 52 columns wide (partial inner blocks); pivots must equal the oracle's.
    rng = np.random.default_rng(22)
    n = 2100
    a = rng.uniform(-1, 1, (1, n, n))
    b = rng.standard_normal((1, n, 1))
    lu, x, ipiv, info = gesv_host(a, b)
    assert info[0] == 0
    xo, luo, steps = oracle.solve(a[0], b[0], with_factors=True)
    np.testing.assert_array_equal(ipiv[0], steps)
    assert np.abs(lu[0] - luo).max() <= 1e-10 * np.abs(luo).max()
    assert np.abs(x[0] - xo).max() <= 1e-13 * np.linalg.cond(a[0], 1) * np.abs(xo).max()


@pytest.mark.parametrize("n", [100, 128, 200, 256])
def test_many_systems_thin_leaves(hip, n):
    # 1100 systems (at least four per CU): the panels run one or two waves per
    # matrix. General matrices, so rows move at nearly every step; pivots of a
    # sample against the oracle, every solution by its backward error.
    rng = np.random.default_rng(40 + n)
    count = 1100
    a = rng.uniform(-1, 1, (count, n, n))
    b = rng.standard_normal((count, n, 2))
    lu, x, ipiv, info = gesv_host(a, b)
    assert np.all(info == 0)
    r = b - a @ x
    eta = np.abs(r).max(axis=1) / (np.abs(a).sum(axis=2).max(axis=1)[:, None] * np.abs(x).max(axis=1)
                                   + np.abs(b).max(axis=1))
    assert eta.max() < 1e-14, eta.max()
    for i in range(0, count, 137):
        xo, luo, steps = oracle.solve(a[i], b[i], with_factors=True)
        if pivot_gaps(a[i]) > 1e-6:
            np.testing.assert_array_equal(ipiv[i], steps, err_msg=f"n={n} system {i}")
            assert np.abs(lu[i] - luo).max() <= 1e-11 * np.abs(luo).max(), (n, i)
        assert np.abs(x[i] - xo).max() <= 1e-13 * np.linalg.cond(a[i], 1) * np.abs(xo).max(), (n, i)


def test_singular_exact_zero(hip):
    n = 70
    rng = np.random.default_rng(4)
    a = rng.uniform(-1, 1, (4, n, n))
    a[0, :, 5] = 0.0                    # zero column: zero pivot at step 5
    a[1, 10] = 0.0                      # zero row
    a[2] = 0.0                          # all zero: step 0
    b = rng.standard_normal((4, n, 1))
    _, _, _, info = gesv_host(a, b)
    for i in range(4):
        sing = oracle.lu_decompose(a[i])[4]
        assert info[i] == (0 if sing < 0 else sing + 1), (i, info[i], sing)
    small = np.zeros((2, 8, 8))
    small[0] = np.eye(8)
    small[0, 3, 3] = 0.0
    small[1] = np.eye(8)
    _, _, _, info = gesv_host(small, np.ones((2, 8, 1)))
    assert info.tolist() == [4, 0]


def test_solve_mirror_semantics(hip, monkeypatch):
    # tests/test_batched_linalg.cpp:35-51 stacks, on the GPU hook
    A = np.arange(12.0).reshape(3, 2, 2)
    b = np.arange(1.0, 7.0).reshape(3, 2)
    x = nla.solve(A, b, Backend.Hip)
    assert last_backend() == Backend.Hip and x.shape == (3, 2)
    np.testing.assert_allclose(x, np.linalg.solve(A, b[..., None])[..., 0], rtol=1e-9, atol=1e-11)
    B = np.arange(1.0, 13.0).reshape(3, 2, 2)
    np.testing.assert_allclose(nla.solve(A, B, Backend.Hip), np.linalg.solve(A, B), rtol=1e-9, atol=1e-11)
    # 2-D, float32 preserved, env target
    monkeypatch.setenv("NUMPP_GPU_TARGET", "hip")
    x32 = nla.solve(np.eye(3, dtype=np.float32) * 2, np.ones(3, np.float32))
    assert x32.dtype == np.float32 and np.allclose(x32, 0.5)
    with pytest.raises(nla.LinAlgError):
        nla.solve(np.zeros((2, 3, 3)), np.ones((2, 3)))


def test_device_getrf_getrs(hip):
    rng = np.random.default_rng(9)
    for n in (40, 150):
        a = rng.uniform(-1, 1, (5, n, n))
        b = rng.standard_normal((5, n, 4))
        dA = DeviceArray.from_host(a)
        dP = DeviceArray((5, n), np.int32)
        dI = DeviceArray((5,), np.int32)
        nla.getrf_batched_dev(dA, dP, dI)
        dB = DeviceArray.from_host(b)
        nla.getrs_batched_dev(dA, dP, dB)
        x = dB.to_host()
        assert np.all(dI.to_host() == 0)
        for i in range(5):
            xo = oracle.solve(a[i], b[i])
            assert np.abs(x[i] - xo).max() <= 1e-13 * np.linalg.cond(a[i], 1) * np.abs(xo).max()
