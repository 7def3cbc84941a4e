"""Uneven batches on the GPU: every system against the oracle, bucket padding
leaves results identical to the uniform path, graph replays match eager runs,
and a system's result does not depend on the rest of the batch."""
import numpy as np
import pytest

import oracle
from conftest import cond_matrix
from numpp_amd import vbatch
from numpp_amd.device import DeviceArray

pytestmark = pytest.mark.gpu


def make(seed, count, nlo=8, nhi=256, rmax=16, dominant=True):
    rng = np.random.default_rng(seed)
    ns = rng.integers(nlo, nhi + 1, count)
    rs = rng.integers(1, rmax + 1, count)
    mats, rhss = [], []
    for n, r in zip(ns, rs):
        a = rng.uniform(-1, 1, (n, n))
        if dominant:
            a += n * np.eye(n)
        mats.append(a)
        rhss.append(rng.standard_normal((n, r)))
    return mats, rhss


def test_fp64_uneven_vs_oracle(hip):
    mats, rhss = make(4, 60)
    xs, berr, iters, info = vbatch.solve_uneven(mats, rhss, "fp64")
    assert np.all(info == 0)
    for a, b, x in zip(mats, rhss, xs):
        xo = oracle.solve(a, b)
        assert np.abs(x - xo).max() <= 1e-12 * np.abs(xo).max()


def test_general_matrices_pivoting(hip):
    mats, rhss = make(5, 30, 20, 200, 5, dominant=False)
    xs, _, _, info = vbatch.solve_uneven(mats, rhss, "fp64")
    assert np.all(info == 0)
    for a, b, x in zip(mats, rhss, xs):
        xo = oracle.solve(a, b)
        assert np.abs(x - xo).max() <= 1e-13 * np.linalg.cond(a, 1) * np.abs(xo).max()


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_mixed_uneven(hip, precision):
    rng = np.random.default_rng(6)
    mats, rhss = [], []
    for n in rng.integers(8, 257, 24):
        mats.append(cond_matrix(rng, int(n), 1e4))
        rhss.append(rng.standard_normal((int(n), int(rng.integers(1, 17)))))
    xs, berr, iters, info = vbatch.solve_uneven(mats, rhss, precision)
    assert np.all(info == 0)
    off = 0
    for a, b, x in zip(mats, rhss, xs):
        eta = oracle.backward_error(a, b, x)
        assert np.all(eta <= 1e-12)
        assert np.all(berr[off:off + b.shape[1]] <= 1e-12)
        off += b.shape[1]


def test_graph_replay_matches_eager(hip):
    mats, rhss = make(7, 80)
    A, B, n, nr = vbatch.pack(mats, rhss)
    plan = vbatch.VPlan(n, nr, "fp64")
    dA, dB = DeviceArray.from_host(A), DeviceArray.from_host(B)
    dX, dI = DeviceArray(B.shape), DeviceArray((len(n),), np.int32)
    plan.execute(dA, dB, dX, dinfo=dI, use_graph=False)
    eager = dX.to_host()
    for rep in range(4):  # first graph call runs eagerly, then capture, then replays
        dX.copy_from(np.zeros_like(B))
        plan.execute(dA, dB, dX, dinfo=dI, use_graph=True)
        np.testing.assert_array_equal(dX.to_host(), eager)
        # program's own eager work between replays
        plan.execute(dA, dB, dX, dinfo=dI, use_graph=False)
    plan.close()


def test_alone_equals_in_mix(hip):
    mats, rhss = make(8, 40)
    xs, _, _, _ = vbatch.solve_uneven(mats, rhss, "fp64")
    for i in (0, 17, 39):
        x1, _, _, _ = vbatch.solve_uneven([mats[i]], [rhss[i]], "fp64")
        np.testing.assert_array_equal(x1[0], xs[i])
