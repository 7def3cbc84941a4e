"""ROCm DLPack exchange of device-resident arrays (kDLROCM = 10; NumPP's
dlpack.h:9 lacks it and interop.cpp:35 hard-codes kDLCPU). Checked with torch
on ROCm as the other framework: both directions are zero-copy views (a write on
one side is seen on the other), dtype / shape / device survive, a consumed
export keeps the memory alive until the consumer releases it, and a solve on
an imported tensor writes the framework's memory in place."""
import gc

import numpy as np
import pytest

import oracle
from numpp_amd import linalg as nla
from numpp_amd.device import DeviceArray, kDLROCM

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_export_to_torch(hip):
    a = np.arange(12.0).reshape(3, 4)
    d = DeviceArray.from_host(a)
    assert d.__dlpack_device__() == (kDLROCM, d.device)
    t = torch.from_dlpack(d)
    assert t.is_cuda and t.dtype == torch.float64 and tuple(t.shape) == (3, 4)
    np.testing.assert_array_equal(t.cpu().numpy(), a)
    t.mul_(2.0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.to_host(), 2 * a)  # same memory
    del d
    gc.collect()
    np.testing.assert_array_equal(t.cpu().numpy(), 2 * a)  # kept alive by the consumer


def test_import_from_torch_and_solve_in_place(hip):
    rng = np.random.default_rng(3)
    a = rng.uniform(-1, 1, (4, 80, 80)) + 80 * np.eye(80)
    b = rng.standard_normal((4, 80, 2))
    ta = torch.tensor(a, device="cuda")
    tb = torch.tensor(b, device="cuda")
    torch.cuda.synchronize()
    dA = DeviceArray.from_dlpack(ta)
    dB = DeviceArray.from_dlpack(tb)
    assert dA.shape == (4, 80, 80) and dA.dtype == np.float64
    nla.gesv_batched_dev(dA, dB)
    from numpp_amd import _lib
    _lib.check(_lib.load().npph_device_sync())
    x = tb.cpu().numpy()  # the solve wrote torch's memory
    for i in range(4):
        xo = oracle.solve(a[i], b[i])
        assert np.abs(x[i] - xo).max() <= 1e-12 * np.abs(xo).max()
    dA.free()
    dB.free()
    assert tb.sum().item() == pytest.approx(float(x.sum()))


def test_rejects_non_contiguous_and_cpu(hip):
    t = torch.zeros((8, 8), device="cuda", dtype=torch.float64).t()
    with pytest.raises(ValueError):
        DeviceArray.from_dlpack(t)
    with pytest.raises(TypeError):
        DeviceArray.from_dlpack(torch.zeros(4, dtype=torch.float64))
