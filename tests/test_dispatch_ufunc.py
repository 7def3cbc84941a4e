"""Dispatch rules of the ufunc / reduction mirror (ufunc.cpp:519-545, 606-617)
and of the LAPACK probe and tensorsolve in linalg, checked on the host: without
a device (or below the size threshold, or for a declined case) NumPP's CPU
kernel runs and last_backend() is Cpu."""
import numpy as np
import pytest

import numpp_amd
from numpp_amd import Backend, CpuPathSelected, last_backend, ufunc
from numpp_amd import linalg as nla


def no_device():
    return not numpp_amd.available()


def test_below_threshold_uses_cpu(monkeypatch):
    monkeypatch.setenv("NUMPP_GPU_MIN", "1000")
    a = np.ones(10)
    assert np.array_equal(ufunc.add(a, a, cpu_kernel=lambda: a + a), 2 * a)
    assert last_backend() == Backend.Cpu
    with pytest.raises(CpuPathSelected):
        ufunc.sqrt(a)


def test_shape_mismatch_and_integer_dtypes_use_cpu(monkeypatch):
    monkeypatch.setenv("NUMPP_GPU_MIN", "0")
    a, b = np.ones(4), np.ones((2, 4))
    assert ufunc.multiply(a, b, cpu_kernel=lambda: a * b).shape == (2, 4)
    assert last_backend() == Backend.Cpu
    i = np.arange(6)
    r = ufunc.add(i, i, cpu_kernel=lambda: i + i)
    assert r.dtype == i.dtype and last_backend() == Backend.Cpu
    assert ufunc.sum(i, cpu_kernel=lambda: i.sum()) == 15


def test_reduce_with_axis_uses_cpu(monkeypatch):
    monkeypatch.setenv("NUMPP_GPU_MIN", "0")
    a = np.ones((3, 4))
    r = ufunc.sum(a, axis=0, cpu_kernel=lambda: a.sum(axis=0))
    assert r.shape == (4,) and last_backend() == Backend.Cpu


@pytest.mark.skipif(not no_device(), reason="checks the no-device path")
def test_no_device_everything_on_cpu(monkeypatch):
    monkeypatch.setenv("NUMPP_GPU_MIN", "0")
    a = np.linspace(0, 1, 100)
    assert np.array_equal(ufunc.divide(a, a + 1, cpu_kernel=lambda: a / (a + 1)), a / (a + 1))
    assert last_backend() == Backend.Cpu
    monkeypatch.setenv("NUMPP_LAPACK_BACKEND", "hip")
    assert nla.lapack_vtable() is None


def test_divide_dtype_rule():
    from numpp_amd.ufunc import _arith_out, GPU_DIV, GPU_ADD
    i = np.arange(3)
    assert _arith_out(GPU_DIV, i, i) == np.float64
    assert _arith_out(GPU_ADD, i, i) == i.dtype
    f = np.ones(2, np.float32)
    assert _arith_out(GPU_DIV, f, f) == np.float32
    with pytest.raises(TypeError):
        _arith_out(1, np.ones(2, bool), np.ones(2, bool))


def test_tensorsolve_shapes_on_cpu(golden):
    g = golden("tensor_cases.npz")
    x = nla.tensorsolve(g["M"], g["b"], cpu_kernel=np.linalg.solve)
    np.testing.assert_allclose(x, g["x"], rtol=1e-12)
    with pytest.raises(nla.LinAlgError):
        nla.tensorsolve(np.ones((2, 3)), np.ones((2, 3, 4)))
    with pytest.raises(nla.LinAlgError):
        nla.tensorsolve(np.ones((2, 2, 3)), np.ones((2, 2)))
    with pytest.raises(ValueError):
        nla.tensorinv(g["M"], 0)
