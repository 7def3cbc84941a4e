"""det / slogdet / inv on the GPU against the oracle's restatement of
linalg.cpp:105-115 and 393-439 (and batched.cpp:75-76, 102-108).

Tolerances (stated): det relative error <= 1e-13 * n * cond_1 (the LU's
backward error moves det by at most about n u cond relative), logabsdet
absolute error <= the same bound, sign exact; inv within 1e-12 * cond_1
relative (max norm) and A inv(A) = I to 1e-12 * cond_1. Exact results where
the arithmetic is exact: the reference literals (det of arange(12)
.reshape(3,2,2) is -2 per matrix, det [[4,2,1],[2,5,3],[1,3,6]] = 67),
permutation matrices (det = sign of the permutation, exactly), 0x0 (det 1),
and exact-zero singular inputs (det 0, slogdet (0, -inf), inv raises).
"""
import numpy as np
import pytest

import oracle
from numpp_amd import Backend, last_backend
from numpp_amd import linalg as nla
from numpp_amd.device import DeviceArray

pytestmark = pytest.mark.gpu


def check_stack(a):
    d = nla.det(a, Backend.Hip)
    assert last_backend() == Backend.Hip
    sg, la = nla.slogdet(a, Backend.Hip)
    ai = nla.inv(a, Backend.Hip)
    a3 = a.reshape(-1, a.shape[-1], a.shape[-1])
    n = a.shape[-1]
    d, sg, la, ai = np.ravel(d), np.ravel(sg), np.ravel(la), ai.reshape(a3.shape)
    for i in range(a3.shape[0]):
        c = np.linalg.cond(a3[i], 1)
        do = oracle.det(a3[i])
        so, lo = oracle.slogdet(a3[i])
        tol = 1e-13 * max(n, 1) * c
        assert abs(d[i] - do) <= tol * abs(do), (i, d[i], do)
        assert sg[i] == so and abs(la[i] - lo) <= tol, (i, sg[i], so, la[i], lo)
        io = oracle.inv(a3[i])
        assert np.abs(ai[i] - io).max() <= 1e-12 * c * np.abs(io).max(), i
        assert np.abs(a3[i] @ ai[i] - np.eye(n)).max() <= 1e-12 * c, i


@pytest.mark.parametrize("n", [1, 2, 7, 33, 64, 65, 130, 257])
def test_det_inv_stacks(hip, n):
    rng = np.random.default_rng(n)
    a = rng.uniform(-1, 1, (5, n, n)) + np.diag(rng.uniform(0.5, 1.5, n))
    a = a[:, rng.permutation(n)]  # rows shuffled: pivoting swaps rows
    check_stack(a)


def test_det_inv_lead_shape_and_goldens(hip, golden):
    g = golden("det_cases.npz")
    for n in g["sizes"]:
        a = g[f"A_{n}"]
        d = nla.det(a, Backend.Hip)
        assert abs(d - g[f"det_{n}"]) <= 1e-12 * n * abs(g[f"det_{n}"])
        s, la = nla.slogdet(a, Backend.Hip)
        assert s == g[f"sign_{n}"] and abs(la - g[f"logabsdet_{n}"]) <= 1e-12 * n * max(1.0, abs(la))
        np.testing.assert_allclose(nla.inv(a, Backend.Hip), g[f"inv_{n}"], rtol=0,
                                   atol=1e-11 * np.abs(g[f"inv_{n}"]).max())
    rng = np.random.default_rng(3)
    a = rng.uniform(-1, 1, (2, 3, 20, 20)) + 3 * np.eye(20)
    assert nla.det(a, Backend.Hip).shape == (2, 3)
    assert nla.slogdet(a, Backend.Hip)[1].shape == (2, 3)
    assert nla.inv(a, Backend.Hip).shape == (2, 3, 20, 20)
    check_stack(a)


def test_reference_literals(hip):
    # test_batched_linalg.cpp:21-32, 66-74: arange(12.).reshape(3,2,2), det -2 each
    A = np.arange(12.0).reshape(3, 2, 2)
    np.testing.assert_allclose(nla.det(A, Backend.Hip), [-2.0] * 3, rtol=1e-14)
    s, la = nla.slogdet(A, Backend.Hip)
    assert np.all(s == -1.0)
    np.testing.assert_allclose(la, [np.log(2.0)] * 3, rtol=1e-14)
    np.testing.assert_allclose(nla.inv(A, Backend.Hip), np.linalg.inv(A), rtol=1e-9, atol=1e-11)
    # test_linalg.cpp:40-54: M = [[4,2,1],[2,5,3],[1,3,6]], det 67
    M = np.array([[4.0, 2, 1], [2, 5, 3], [1, 3, 6]])
    assert abs(nla.det(M, Backend.Hip) - 67.0) <= 1e-13 * 67
    np.testing.assert_allclose(nla.inv(M, Backend.Hip) @ M, np.eye(3), atol=1e-14)
    # test_linalg_mined.cpp:82-100: 0x0 and 1x1
    assert nla.det(np.empty((0, 0)), Backend.Hip) == 1.0
    assert nla.slogdet(np.empty((0, 0)), Backend.Hip) == (1.0, 0.0)
    assert nla.det(np.array([[5.0]]), Backend.Hip) == 5.0
    assert nla.inv(np.array([[5.0]]), Backend.Hip)[0, 0] == 0.2


def test_permutation_sign_exact(hip):
    rng = np.random.default_rng(9)
    for n in (3, 40, 100):
        p = rng.permutation(n)
        inversions = sum(1 for i in range(n) for j in range(i + 1, n) if p[i] > p[j])
        assert nla.det(np.eye(n)[p], Backend.Hip) == (-1.0) ** inversions
        assert nla.slogdet(np.eye(n)[p], Backend.Hip) == ((-1.0) ** inversions, 0.0)


def test_singular(hip):
    # test_linalg.cpp:56-59 (zeros) and an exact rank deficiency at a late step
    z = np.zeros((2, 2))
    assert nla.det(z, Backend.Hip) == 0.0
    assert nla.slogdet(z, Backend.Hip) == (0.0, -np.inf)
    with pytest.raises(nla.LinAlgError):
        nla.inv(z, Backend.Hip)
    for n in (6, 90):
        a = np.eye(n)
        a[n - 2, n - 2] = 0.0
        st = np.stack([np.eye(n) * 2.0, a])
        np.testing.assert_array_equal(nla.det(st, Backend.Hip), [2.0 ** n, 0.0])
        s, la = nla.slogdet(st, Backend.Hip)
        assert s.tolist() == [1.0, 0.0] and la[1] == -np.inf
        with pytest.raises(nla.LinAlgError):
            nla.inv(st, Backend.Hip)


def test_float32_in_float32_out(hip):
    rng = np.random.default_rng(4)
    a = (rng.uniform(-1, 1, (3, 16, 16)) + 4 * np.eye(16)).astype(np.float32)
    d = nla.det(a, Backend.Hip)
    assert d.dtype == np.float32
    np.testing.assert_allclose(d, [np.float32(oracle.det(a[i].astype(np.float64))) for i in range(3)], rtol=1e-6)
    assert nla.inv(a, Backend.Hip).dtype == np.float32


def test_device_resident(hip):
    rng = np.random.default_rng(5)
    a = rng.uniform(-1, 1, (4, 100, 100)) + 5 * np.eye(100)
    dA = DeviceArray.from_host(a)
    dD, dS, dL = DeviceArray((4,)), DeviceArray((4,)), DeviceArray((4,))
    dI = DeviceArray((4,), np.int32)
    nla.det_batched_dev(dA, dD, dS, dL, dI)
    assert np.all(dI.to_host() == 0)
    for i in range(4):
        assert abs(dD.to_host()[i] - oracle.det(a[i])) <= 1e-12 * abs(oracle.det(a[i]))
    dA = DeviceArray.from_host(a)
    dX = DeviceArray((4, 100, 100))
    nla.inv_batched_dev(dA, dX, dI)
    np.testing.assert_allclose(dX.to_host() @ a, np.broadcast_to(np.eye(100), a.shape), atol=1e-12)
