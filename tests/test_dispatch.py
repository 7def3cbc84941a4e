"""Dispatcher semantics mirrored from NumPP (backend.cpp:121-177,
linalg.cpp:368-385, batched.cpp:90-100), exercised on the CPU: the CPU branch
runs the oracle passed as NumPP's CPU kernel."""
import numpy as np
import pytest

import oracle
from numpp_amd import Backend, CpuPathSelected, last_backend, matmul
from numpp_amd import linalg as nla
from numpp_amd.backend import env_target, gpu_available


def cpu_matmul(a, b):
    return oracle.matmul(a, b)


def cpu_solve(a, b):
    if a.ndim > 2:
        bvec = b.ndim == a.ndim - 1
        x, info = oracle.batched_solve(a.astype(np.float64), b.astype(np.float64))
        if np.any(info >= 0):
            raise nla.LinAlgError("Singular matrix")
        return x
    return oracle.solve(a.astype(np.float64), b.astype(np.float64))


def test_matmul_cpu_dispatch_literal():
    a = np.arange(1.0, 7.0).reshape(2, 3)
    b = np.arange(1.0, 7.0).reshape(3, 2)
    c = matmul(a, b, Backend.Cpu, cpu_kernel=cpu_matmul)
    assert last_backend() == Backend.Cpu
    assert c[0, 0] == 22.0 and c[1, 1] == 64.0


def test_matmul_shape_errors():
    with pytest.raises(ValueError):
        matmul(np.ones(3), np.ones((3, 3)))
    with pytest.raises(ValueError):
        matmul(np.ones((2, 3)), np.ones((2, 3)))


def test_auto_without_cpu_kernel_raises():
    with pytest.raises(CpuPathSelected):
        matmul(np.ones((2, 2)), np.ones((2, 2)))


def test_env_target(monkeypatch):
    monkeypatch.setenv("NUMPP_GPU_TARGET", "cpu")
    assert env_target() == Backend.Cpu
    matmul(np.ones((2, 2)), np.ones((2, 2)), cpu_kernel=cpu_matmul)
    assert last_backend() == Backend.Cpu
    monkeypatch.setenv("NUMPP_GPU_TARGET", "hip")
    assert env_target() == Backend.Hip
    monkeypatch.setenv("NUMPP_GPU_TARGET", "bogus")
    assert env_target() == Backend.Auto


def test_explicit_unavailable_backend_errors(monkeypatch):
    # tests/test_backend.cpp:75-88: explicit GPU not available -> not_implemented_error
    a = np.ones((2, 2))
    for b in (Backend.Cuda, Backend.OpenCL):
        with pytest.raises(NotImplementedError):
            matmul(a, a, b)
    with pytest.raises(NotImplementedError):
        matmul(a, a, Backend.Blas, cpu_kernel=cpu_matmul)
    if not gpu_available(Backend.Hip):
        with pytest.raises(NotImplementedError):
            matmul(a, a, Backend.Hip)
        monkeypatch.setenv("NUMPP_GPU_TARGET", "hip")
        with pytest.raises(NotImplementedError):
            nla.solve(np.eye(2), np.ones(2))


def test_solve_cpu_path_semantics():
    A = np.arange(12.0).reshape(3, 2, 2)
    b = np.arange(1.0, 7.0).reshape(3, 2)
    x = nla.solve(A, b, Backend.Cpu, cpu_kernel=cpu_solve)
    assert x.shape == (3, 2)
    np.testing.assert_allclose(np.einsum("kij,kj->ki", A, x), b, atol=1e-12)
    assert last_backend() == Backend.Cpu
    with pytest.raises(nla.LinAlgError):
        nla.solve(np.ones((2, 3)), np.ones(2), Backend.Cpu, cpu_kernel=cpu_solve)
    # float32 in -> float32 out (tests/test_linalg_mined.cpp:191-196)
    x32 = nla.solve(np.eye(3, dtype=np.float32), np.ones(3, np.float32), Backend.Cpu, cpu_kernel=cpu_solve)
    assert x32.dtype == np.float32
    with pytest.raises(ValueError):
        nla.solve(np.eye(3), np.ones(4), Backend.Cpu, cpu_kernel=cpu_solve)
