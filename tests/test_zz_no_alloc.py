"""The allocation-free stacked solve (npph_gesv_batched_dev_noalloc) through
the C ABI against npph_gesv_batched_dev, to the same bar as the plain solve:
bitwise the same factors, pivots, info and solutions (both run the same
kernels), on two calls in a row with the second smaller; pivots equal to the
oracle's and x within 1e-12 relative (well-conditioned systems); a call whose
scratch would exceed the fixed maximum fails before touching its inputs."""
import numpy as np
import pytest

import oracle
from numpp_amd import _lib
from numpp_amd.device import DeviceArray

pytestmark = pytest.mark.gpu


def run(fn, a, b, with_piv=True):
    count, n, _ = a.shape
    nrhs = b.shape[2]
    dA, dB = DeviceArray.from_host(a), DeviceArray.from_host(b)
    dP = DeviceArray((count, n), np.int32) if with_piv else None
    dI = DeviceArray((count,), np.int32) if with_piv else None
    _lib.check(fn(count, n, nrhs, dA.ptr, dB.ptr, dP.ptr if dP else None, dI.ptr if dI else None, None))
    _lib.check(_lib.load().npph_device_sync())
    return dA.to_host(), dB.to_host(), (dP.to_host() if dP else None), (dI.to_host() if dI else None)


@pytest.mark.parametrize("shapes", [[(8, 300, 4), (3, 130, 2)], [(64, 64, 1), (5, 33, 3)],
                                    [(2, 2048, 8), (1, 700, 1)]])
def test_noalloc_matches_gesv(hip, shapes):
    lib = _lib.load()
    for count, n, nrhs in shapes:
        rng = np.random.default_rng(count * 1000 + n)
        a = rng.uniform(-1, 1, (count, n, n)) + n * np.eye(n)
        a = a[:, rng.permutation(n)]
        b = rng.standard_normal((count, n, nrhs))
        ref = run(lib.npph_gesv_batched_dev, a, b)
        got = run(lib.npph_gesv_batched_dev_noalloc, a, b)
        for r, g in zip(ref, got):
            np.testing.assert_array_equal(g, r)
        _, x_np, _, _ = run(lib.npph_gesv_batched_dev_noalloc, a, b, with_piv=False)
        np.testing.assert_array_equal(x_np, ref[1])
        assert np.all(got[3] == 0)
        for i in range(min(count, 3)):
            xo, _, steps = oracle.solve(a[i], b[i], with_factors=True)
            np.testing.assert_array_equal(got[2][i], steps)
            assert np.abs(got[1][i] - xo).max() <= 1e-12 * np.abs(xo).max()


def test_noalloc_rejects_oversize(hip):
    lib = _lib.load()
    count, n, nrhs = 8, 2048, 8
    rng = np.random.default_rng(1)
    a = rng.uniform(-1, 1, (count, n, n))
    dA = DeviceArray.from_host(a)
    dB = DeviceArray((count, n, nrhs))
    st = lib.npph_gesv_batched_dev_noalloc(count, n, nrhs, dA.ptr, dB.ptr, None, None, None)
    assert st == _lib.NPPH_ERR_INVALID
    assert "maximum" in _lib.last_error()
    np.testing.assert_array_equal(dA.to_host(), a)
