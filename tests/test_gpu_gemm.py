"""GPU GEMM parity: the MFMA kernels against the oracle's cpu_matmul restatement
(backend.cpp:55-84) with the componentwise bound |dC| <= 2 gamma_k (|A||B|),
gamma_k = k u / (1 - k u), and the reference's own rtol = atol = 1e-10 device
check (tests/test_gpu.cpp:67-69)."""
import numpy as np
import pytest

import oracle
from numpp_amd import Backend, last_backend, matmul
from numpp_amd import _lib
from numpp_amd.device import DeviceArray

pytestmark = pytest.mark.gpu


def gamma(k, dt):
    u = np.finfo(dt).eps / 2
    return k * u / (1 - k * u)


def check_bound(c, a, b, ref, dt):
    k = a.shape[1]
    mag = np.abs(a).astype(np.float64) @ np.abs(b).astype(np.float64)
    bound = 2 * gamma(k, dt) * mag + np.finfo(dt).tiny
    err = np.abs(c.astype(np.float64) - ref.astype(np.float64))
    assert np.all(err <= bound), f"max excess {(err - bound).max()}"


def test_literal_and_reference_golden(hip, golden):
    a = np.arange(1.0, 7.0).reshape(2, 3)
    b = np.arange(1.0, 7.0).reshape(3, 2)
    c = matmul(a, b, Backend.Hip)
    assert last_backend() == Backend.Hip
    assert c.tolist() == [[22.0, 28.0], [49.0, 64.0]]
    a = np.arange(1, 21, dtype=np.float64).reshape(4, 5)
    b = np.arange(2, 17, dtype=np.float64).reshape(5, 3)
    np.testing.assert_array_equal(matmul(a, b, Backend.Hip), golden("numpp_g15653376641833044941.npy"))


def test_reference_device_gemm_case(hip):
    # tests/test_gpu.cpp:56-79: C(40x30) = A(40x50) @ B(50x30)
    A = np.arange(0.0, 2000.0).reshape(40, 50)
    B = np.arange(0.0, 1500.0).reshape(50, 30) * 0.001
    cpu = oracle.matmul(A, B)
    dev = matmul(A, B, Backend.Hip)
    np.testing.assert_allclose(dev, cpu, rtol=1e-10, atol=1e-10)


def test_numpy_goldens_f64_f32(hip, golden):
    g = golden("matmul_cases.npz")
    for t in range(int(g["count"][0])):
        a, b = g[f"A64_{t}"], g[f"B64_{t}"]
        c = matmul(a, b, Backend.Hip)
        check_bound(c, a, b, oracle.matmul(a, b), np.float64)
        a32, b32 = g[f"A32_{t}"], g[f"B32_{t}"]
        c32 = matmul(a32, b32, Backend.Hip)
        assert c32.dtype == np.float32
        check_bound(c32, a32, b32, oracle.matmul(a32, b32), np.float32)


@pytest.mark.parametrize("m,k,n", [(1, 1, 1), (127, 129, 131), (128, 16, 128), (256, 200, 130), (513, 64, 257),
                                   (1000, 333, 77), (64, 1, 64), (3, 1024, 5)])
def test_edge_shapes(hip, m, k, n):
    rng = np.random.default_rng(m * 7 + k * 3 + n)
    a = rng.standard_normal((m, k))
    b = rng.standard_normal((k, n))
    check_bound(matmul(a, b, Backend.Hip), a, b, oracle.matmul(a, b), np.float64)


def test_1024_full_against_oracle(hip):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((1024, 1024))
    b = rng.standard_normal((1024, 1024))
    c = matmul(a, b, Backend.Hip)
    check_bound(c, a, b, oracle.matmul(a, b), np.float64)
    np.testing.assert_allclose(c, oracle.matmul(a, b), rtol=1e-10, atol=1e-10)


def test_8192_sampled_dot2(hip):
    # The GEMM workload: A, B standard normal from default_rng(1), A drawn first.
    rng = np.random.default_rng(1)
    n = 8192
    a = rng.standard_normal((n, n))
    b = rng.standard_normal((n, n))
    dA, dB = DeviceArray.from_host(a), DeviceArray.from_host(b)
    dC = DeviceArray((n, n))
    _lib.check(_lib.load().npph_dgemm_dev(n, n, n, 1.0, dA.ptr, n, 0, dB.ptr, n, 0, 0.0, dC.ptr, n, 0, 1, None))
    c = dC.to_host()
    srng = np.random.default_rng(99)
    rows = srng.integers(0, n, 256)
    cols = srng.integers(0, n, 256)
    rows = np.concatenate([rows, [0, n - 1, 0, n - 1]])
    cols = np.concatenate([cols, [0, n - 1, n - 1, 0]])
    ref, mag = oracle.dot2_entries(a, b, rows, cols)
    err = np.abs(c[rows, cols] - ref)
    assert np.all(err <= 2 * gamma(n, np.float64) * mag)
    # a checksum of checksums: e^T C e = (e^T A)(B e)
    lhs = c.sum()
    rhs = a.sum(0) @ b.sum(1)
    assert abs(lhs - rhs) <= 1e-9 * (np.abs(a).sum(0) @ np.abs(b).sum(1))


def test_batched_strided_dev(hip):
    rng = np.random.default_rng(3)
    bs, m, k, n = 5, 70, 65, 90
    a = rng.standard_normal((bs, m, k))
    b = rng.standard_normal((bs, k, n))
    c0 = rng.standard_normal((bs, m, n))
    dA, dB, dC = DeviceArray.from_host(a), DeviceArray.from_host(b), DeviceArray.from_host(c0)
    _lib.check(_lib.load().npph_dgemm_dev(m, n, k, -1.0, dA.ptr, k, m * k, dB.ptr, n, k * n, 1.0, dC.ptr, n,
                                          m * n, bs, None))
    c = dC.to_host()
    for i in range(bs):
        ref = c0[i] - oracle.matmul(a[i], b[i])
        assert np.abs(c[i] - ref).max() <= 4 * gamma(k, np.float64) * (np.abs(a[i]) @ np.abs(b[i]) + np.abs(c0[i])).max()


@pytest.mark.parametrize("terms,bits", [(1, 6), (3, 14), (6, 17)])
@pytest.mark.parametrize("m,k,n", [(300, 64, 260), (257, 1000, 129), (128, 33, 384)])
def test_bf16_split_update(hip, terms, bits, m, k, n):
    rng = np.random.default_rng(terms + m)
    a = rng.standard_normal((m, k)).astype(np.float32)
    b = rng.standard_normal((k, n)).astype(np.float32) * 1e-3
    c0 = rng.standard_normal((m, n)).astype(np.float32)
    dA, dB, dC = DeviceArray.from_host(a), DeviceArray.from_host(b), DeviceArray.from_host(c0)
    _lib.check(_lib.load().npph_bf16split_update_dev(m, n, k, dA.ptr, k, dB.ptr, n, dC.ptr, n, terms, None))
    c = dC.to_host().astype(np.float64)
    ref = c0.astype(np.float64) - a.astype(np.float64) @ b.astype(np.float64)
    mag = np.abs(a).astype(np.float64) @ np.abs(b).astype(np.float64)
    err = np.abs(c - ref) - np.finfo(np.float32).eps * np.abs(ref)  # C's own FP32 rounding
    assert np.all(err <= 4 * 2.0 ** -bits * mag)


def test_bf16_split_exact_on_bf16_values(hip):
    # Inputs exactly representable in BF16 with small integer products: every
    # term count must give the exact result (checks the MFMA operand layout).
    rng = np.random.default_rng(11)
    m, k, n = 130, 64, 140
    a = rng.integers(-4, 5, (m, k)).astype(np.float32)
    b = rng.integers(-4, 5, (k, n)).astype(np.float32)
    for terms in (1, 3, 6):
        dA, dB = DeviceArray.from_host(a), DeviceArray.from_host(b)
        dC = DeviceArray.from_host(np.zeros((m, n), np.float32))
        _lib.check(_lib.load().npph_bf16split_update_dev(m, n, k, dA.ptr, k, dB.ptr, n, dC.ptr, n, terms, None))
        np.testing.assert_array_equal(dC.to_host(), -(a @ b))
