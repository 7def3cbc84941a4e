"""ufunc slots on the GPU against the oracle (NumPP's CPU kernels restated).

add / subtract / multiply / divide / negative / sqrt: bit-exact (IEEE
operations in the dtype; NaN positions equal, NaN payload and sign not
compared), float32 and float64, special values (signed zeros, infinities,
NaN, subnormals, extremes), sizes that leave a vector tail, misaligned device
pointers. sum / prod: |gpu - cpu| <= 2 gamma_n sum|x| (gamma_n = n u / (1 - n u),
the sequential loop's and the tree's error bounds added), exact on integers.
exp and complex input are declined (slot false -> CPU), as in the reference's
CUDA and OpenCL providers.
"""
import ctypes

import numpy as np
import pytest

import oracle
from numpp_amd import Backend, _lib, last_backend, ufunc, vtable
from numpp_amd.device import DeviceArray

pytestmark = pytest.mark.gpu

BIN = {"add": ufunc.add, "subtract": ufunc.subtract, "multiply": ufunc.multiply, "divide": ufunc.divide}
UN = {"negative": ufunc.negative, "sqrt": ufunc.sqrt}


def inputs(dt, n, seed):
    rng = np.random.default_rng(seed)
    f = np.finfo(dt)
    s = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, f.tiny, -f.tiny, f.smallest_subnormal,
                  f.max, -f.max, 3.0, 1e-3, 7.5], dtype=dt)
    a = np.concatenate([np.repeat(s, s.size), (rng.standard_normal(n) * 10).astype(dt)])
    b = np.concatenate([np.tile(s, s.size), (rng.standard_normal(n) * 10).astype(dt)])
    return a, b


def same_ieee(got, want):
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    u = np.uint32 if got.dtype == np.float32 else np.uint64
    return np.array_equal(got[ok].view(u), want[ok].view(u))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("n", [70000, 100003])
def test_elementwise_exact(hip, monkeypatch, dt, n):
    monkeypatch.setenv("NUMPP_GPU_MIN", "0")
    a, b = inputs(dt, n, n)
    for name, fn in BIN.items():
        got = fn(a, b)
        assert last_backend() == Backend.Device and got.dtype == dt
        assert same_ieee(got, oracle.elementwise(name, a, b)), name
    for name, fn in UN.items():
        got = fn(a)
        assert last_backend() == Backend.Device
        assert same_ieee(got, oracle.elementwise(name, a)), name


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_reductions(hip, monkeypatch, dt):
    monkeypatch.setenv("NUMPP_GPU_MIN", "0")
    rng = np.random.default_rng(5)
    u = np.finfo(dt).eps / 2
    for n in (1, 257, 65536, 1000003):
        x = rng.standard_normal(n).astype(dt)
        s = ufunc.sum(x)
        assert last_backend() == Backend.Device and s.dtype == dt
        g = 2 * n * u / (1 - n * u)
        assert abs(float(s) - float(oracle.reduce("sum", x))) <= 2 * g * float(np.abs(x).astype(np.float64).sum())
        p = rng.uniform(0.999, 1.001, n).astype(dt)
        pr = ufunc.prod(p)
        assert abs(float(pr) - float(oracle.reduce("prod", p))) <= 2 * g * abs(float(oracle.reduce("prod", p)))
    i = np.arange(1, 4097, dtype=dt)
    assert ufunc.sum(i) == dt(4096 * 4097 / 2)
    assert ufunc.sum(np.array([], dt)) == 0 and ufunc.prod(np.array([], dt)) == 1


def test_declined_cases(hip, monkeypatch):
    monkeypatch.setenv("NUMPP_GPU_MIN", "0")
    a = np.linspace(0.1, 1, 1000)
    assert np.array_equal(ufunc.exp(a, cpu_kernel=lambda: np.exp(a)), np.exp(a))
    assert last_backend() == Backend.Cpu
    c = a.astype(np.complex128)
    ufunc.add(c, c, cpu_kernel=lambda: c + c)
    assert last_backend() == Backend.Cpu
    vt = vtable.get()
    assert not vt.elementwise_unary(_lib.UN_EXP, _lib.DT_F64, 4, a.ctypes.data, a.ctypes.data)


def test_device_resident_misaligned(hip):
    lib = _lib.load()
    n = 5001
    a, b = inputs(np.float64, n, 3)
    m = a.size
    dA = DeviceArray.from_host(np.concatenate([[0.0], a]))
    dB = DeviceArray.from_host(np.concatenate([[0.0], b]))
    dO = DeviceArray((m + 1,))
    off = ctypes.c_void_p(dA.ptr.value + 8), ctypes.c_void_p(dB.ptr.value + 8), ctypes.c_void_p(dO.ptr.value + 8)
    _lib.check(lib.npph_elementwise_binary_dev(_lib.EW_DIV, _lib.DT_F64, m, off[0], off[1], off[2], None))
    got = dO.to_host()[1:]
    with np.errstate(all="ignore"):
        assert same_ieee(got, a / b)
    dS = DeviceArray((1,))
    _lib.check(lib.npph_reduce_dev(_lib.RED_SUM, _lib.DT_F64, n, ctypes.c_void_p(dA.ptr.value + 8 * (m - n + 1)),
                                   dS.ptr, None))
    x = a[m - n:]
    assert abs(dS.to_host()[0] - oracle.reduce("sum", x)) <= 4 * n * 1.2e-16 * np.abs(x).sum()
