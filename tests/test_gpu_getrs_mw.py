"""Triangular sweeps with many workgroups per system (stacks of few systems of
order >= 512): X against the oracle's lu_solve restatement (linalg.cpp:73-90).

Tolerance (stated): ||x_gpu - x_cpu||_inf <= 1e-13 * kappa_1 * ||x_cpu||_inf per
system; pivots bit-exact where the oracle certifies them; systems with an
exact-zero pivot are skipped (info = k + 1) without disturbing the others.
Orders that are not multiples of 64 exercise the partial last block, and
right-hand-side counts of 1..33 every column-chunk width (4, 8, 16) and
several chunks per system.
"""
import numpy as np
import pytest

import oracle
from numpp_amd import _lib
from numpp_amd import linalg as nla

pytestmark = pytest.mark.gpu


def gesv_host(a, b):
    a = np.ascontiguousarray(a, np.float64).copy()
    b = np.ascontiguousarray(b, np.float64).copy()
    count, n = a.shape[0], a.shape[1]
    ipiv = np.zeros((count, n), np.int32)
    info = np.zeros(count, np.int32)
    _lib.check(_lib.load().npph_gesv_batched(_lib.DT_F64, count, n, b.shape[2], a.ctypes.data, b.ctypes.data,
                                             ipiv.ctypes.data, info.ctypes.data))
    return a, b, ipiv, info


@pytest.mark.parametrize("n,nrhs,count", [(512, 1, 2), (513, 5, 1), (640, 8, 2), (700, 9, 1), (1000, 16, 1),
                                          (577, 17, 1), (1100, 33, 1), (2048, 4, 1)])
def test_plain_sweeps_vs_oracle(hip, n, nrhs, count):
    rng = np.random.default_rng(n * 7 + nrhs)
    a = rng.uniform(-1, 1, (count, n, n))
    b = rng.standard_normal((count, n, nrhs))
    lu, x, ipiv, info = gesv_host(a, b)
    assert np.all(info == 0)
    for i in range(count):
        xo, luo, steps = oracle.solve(a[i], b[i], with_factors=True)
        kappa = np.linalg.cond(a[i], 1)
        assert np.abs(x[i] - xo).max() <= 1e-13 * kappa * np.abs(xo).max(), (n, nrhs, i)


def test_singular_member_is_skipped(hip):
    rng = np.random.default_rng(5)
    n, nrhs = 600, 3
    a = rng.uniform(-1, 1, (3, n, n))
    a[1][:, 10] = 0.0  # exact-zero column: the elimination meets a zero pivot at step 10
    b = rng.standard_normal((3, n, nrhs))
    lu, x, ipiv, info = gesv_host(a, b)
    assert info[1] == 11 and info[0] == 0 and info[2] == 0
    for i in (0, 2):
        xo = oracle.solve(a[i], b[i])
        assert np.abs(x[i] - xo).max() <= 1e-13 * np.linalg.cond(a[i], 1) * np.abs(xo).max()


@pytest.mark.parametrize("n,nrhs", [(900, 1), (1024, 8), (1300, 20)])
def test_refined_sweeps(hip, n, nrhs):
    # FP32 factors; corrections through the multi-workgroup refinement sweeps
    rng = np.random.default_rng(n + nrhs)
    a = rng.uniform(-1, 1, (2, n, n)) + 0.5 * np.sqrt(n) * np.eye(n)
    b = rng.standard_normal((2, n, nrhs))
    x, berr, iters, info = nla.solve_refined(a, b, precision="fp32")
    assert np.all(info == 0) and np.all(iters >= 1)
    for i in range(2):
        eta = oracle.backward_error(a[i], b[i], x[i])
        assert np.all(eta <= 1e-12), eta
        assert np.all(berr[i] <= 1e-12)


def test_repeated_calls_identical(hip):
    # the sync words are cleared per call: the same inputs give bitwise the same
    # results on consecutive calls (forward and backward flags never go stale)
    rng = np.random.default_rng(9)
    a = rng.uniform(-1, 1, (2, 768, 768))
    b = rng.standard_normal((2, 768, 6))
    r1 = gesv_host(a, b)[1]
    r2 = gesv_host(a, b)[1]
    np.testing.assert_array_equal(r1, r2)
