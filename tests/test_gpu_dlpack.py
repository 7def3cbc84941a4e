"""ROCm DLPack exchange of device-resident arrays (kDLROCM = 10; NumPP's
dlpack.h:9 lacks it and interop.cpp:35 hard-codes kDLCPU). Checked with torch
on ROCm as the other framework: both directions are zero-copy views (a write on
one side is seen on the other), dtype / shape / device survive, a consumed
export keeps the memory alive until the consumer releases it, and a solve on
an imported tensor writes the framework's memory in place.

torch ships its own HIP runtime library; each case runs in one child process
that initialises torch's device first and then this library, as a torch
application using the backend would.
"""
import os
import subprocess
import sys
import textwrap

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

PRELUDE = f"""
import gc, sys
sys.path.insert(0, {ROOT!r})
sys.path.insert(0, {os.path.join(ROOT, 'tests')!r})
import numpy as np
import torch
torch.cuda.init()
import numpp_amd, oracle
from numpp_amd import _lib
from numpp_amd import linalg as nla
from numpp_amd.device import DeviceArray, kDLROCM
numpp_amd.require_device()
"""

CASES = {
    "export": """
a = np.arange(12.0).reshape(3, 4)
d = DeviceArray.from_host(a)
assert d.__dlpack_device__() == (kDLROCM, d.device)
t = torch.from_dlpack(d)
assert t.is_cuda and t.dtype == torch.float64 and tuple(t.shape) == (3, 4)
assert np.array_equal(t.cpu().numpy(), a)
t.mul_(2.0)
torch.cuda.synchronize()
assert np.array_equal(d.to_host(), 2 * a)       # same memory
del d
gc.collect()
assert np.array_equal(t.cpu().numpy(), 2 * a)   # kept alive by the consumer
""",
    "import_solve": """
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
_lib.check(_lib.load().npph_device_sync())
x = tb.cpu().numpy()                              # the solve wrote torch's memory
for i in range(4):
    xo = oracle.solve(a[i], b[i])
    assert np.abs(x[i] - xo).max() <= 1e-12 * np.abs(xo).max()
dA.free(); dB.free()
""",
    "reject": """
t = torch.zeros((8, 8), device="cuda", dtype=torch.float64).t()
try:
    DeviceArray.from_dlpack(t)
    raise SystemExit("non-contiguous accepted")
except ValueError:
    pass
try:
    DeviceArray.from_dlpack(torch.zeros(4, dtype=torch.float64))
    raise SystemExit("cpu tensor accepted")
except TypeError:
    pass
""",
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_dlpack_with_torch(hip, case):
    pytest.importorskip("torch")
    code = PRELUDE + textwrap.dedent(CASES[case]) + "\nprint('ok')\n"
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout[-2000:] + r.stderr[-4000:]
