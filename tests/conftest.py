import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device; runs the HIP library")


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    return load


@pytest.fixture(scope="session")
def hip():
    """The HIP library with a device; fails (never skips) when absent, so a GPU
    run cannot pass without the native code."""
    import numpp_amd
    return numpp_amd.require_device()


def config1_inputs():
    """256 x 64 x 64, A = U[-1,1) + 64 I, b ~ N(0,1), default_rng(0), A drawn first."""
    rng = np.random.default_rng(0)
    a = rng.uniform(-1.0, 1.0, (256, 64, 64)) + 64.0 * np.eye(64)
    b = rng.standard_normal((256, 64))
    return a, b


def cond_matrix(rng, n, cond):
    """A = U diag(s) V^T with Haar U, V (QR of a standard-normal matrix, sign
    fixed by R's diagonal) and s log-spaced from 1 to 1/cond."""
    def haar(k):
        q, r = np.linalg.qr(rng.standard_normal((k, k)))
        return q * np.sign(np.diag(r))
    s = np.logspace(0.0, -np.log10(cond), n)
    return (haar(n) * s) @ haar(n).T
