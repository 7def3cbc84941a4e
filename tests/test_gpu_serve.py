"""Serving loop on the GPU: results against the oracle, bitwise the same alone
and in the mix, resident factors reused, budget admission and eviction."""
import threading

import numpy as np
import pytest

import oracle
from conftest import cond_matrix
from numpp_amd.serve import Server

pytestmark = pytest.mark.gpu


def systems(seed, count, lo=16, hi=300, rmax=8, cond=1e3):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        r = int(rng.integers(1, rmax + 1))
        out.append((cond_matrix(rng, n, cond), rng.standard_normal((n, r))))
    return out


def test_results_and_alone_vs_mix(hip):
    syss = systems(1, 24)
    # repeat a few orders so same-order coalescing happens in the mix
    syss += [(a.copy(), b[:, :1].copy()) for a, b in syss[:6]]
    srv = Server(budget_bytes=8e9, workers=4, max_batch=8, precision="fp32")
    ids = [srv.submit(a, b) for a, b in syss]
    mix = [srv.wait(i) for i in ids]
    st = srv.stats()
    alone = []
    for a, b in syss:
        alone.append(srv.wait(srv.submit(a, b)))
    srv.close()
    for (a, b), m, s in zip(syss, mix, alone):
        assert m["status"] == 0 and s["status"] == 0
        np.testing.assert_array_equal(m["x"], s["x"])
        eta = oracle.backward_error(a, b, m["x"])
        assert np.all(eta <= 1e-12)
        assert np.all(m["berr"] <= 1e-12)
    assert st["completed"] == len(syss)


def test_resident_factor_reuse(hip):
    rng = np.random.default_rng(2)
    a = cond_matrix(rng, 200, 1e4)
    srv = Server(budget_bytes=4e9, workers=2, precision="fp32")
    r0 = srv.wait(srv.submit(a, rng.standard_normal((200, 3)), key=77, keep=True))
    assert r0["status"] == 0 and srv.stats()["resident"] == 1
    for _ in range(3):
        b = rng.standard_normal((200, 5))
        r = srv.wait(srv.submit(None, b, key=77))
        assert r["status"] == 0
        assert np.all(oracle.backward_error(a, b, r["x"]) <= 1e-12)
    # unknown key without A is rejected at submit
    with pytest.raises(Exception):
        srv.submit(None, np.ones(200), key=12345)
    srv.close()


def test_reuse_submitted_while_owner_runs(hip):
    # repeat right-hand sides submitted right behind the request that keeps the
    # factors: they are held until the factors are resident, then served from them
    rng = np.random.default_rng(3)
    n = 700
    a = cond_matrix(rng, n, 1e4)
    bs = [rng.standard_normal((n, 2)) for _ in range(4)]
    srv = Server(budget_bytes=8e9, workers=4, precision="fp32")
    owner = srv.submit(a, bs[0], key=5, keep=True)
    reuse = [srv.submit(None, b, key=5) for b in bs[1:]]
    r0 = srv.wait(owner)
    rs = [srv.wait(i) for i in reuse]
    assert r0["status"] == 0 and all(r["status"] == 0 for r in rs)
    # the same right-hand side through the resident factors after the fact
    late = srv.wait(srv.submit(None, bs[1], key=5))
    np.testing.assert_array_equal(late["x"], rs[0]["x"])
    srv.close()
    for b, r in zip(bs, [r0] + rs):
        assert np.all(oracle.backward_error(a, b, r["x"]) <= 1e-12)


def test_budget_admission_and_eviction(hip):
    rng = np.random.default_rng(3)
    n = 512
    mats = [cond_matrix(rng, n, 1e3) for _ in range(4)]
    per = n * n * 8 * 3.2  # about one request's working bytes
    srv = Server(budget_bytes=per * 2.5, workers=3, precision="fp32")
    ids = [srv.submit(m, rng.standard_normal((n, 2)), key=10 + i, keep=True) for i, m in enumerate(mats)]
    res = [srv.wait(i) for i in ids]
    assert all(r["status"] == 0 for r in res)
    st = srv.stats()
    assert st["peak_bytes"] <= per * 2.5
    assert st["evictions"] >= 1
    # a request larger than the whole budget fails cleanly
    big = srv.submit(np.eye(2048), np.ones(2048))
    assert srv.wait(big)["status"] != 0
    srv.close()


def test_replay_poisson(hip):
    syss = systems(4, 20, hi=400)
    rng = np.random.default_rng(5)
    t = np.cumsum(rng.exponential(0.002, len(syss)))
    srv = Server(budget_bytes=8e9, workers=4, precision="fp32")
    out = srv.replay(t, [a for a, _ in syss], [b for _, b in syss], np.zeros(len(syss), np.uint64),
                     np.zeros(len(syss), np.int32))
    srv.close()
    assert out["status"] == 0 and np.all(out["req_status"] == 0)
    assert np.all(out["latency_ms"] > 0)
    for (a, b), x in zip(syss, out["x"]):
        assert np.all(oracle.backward_error(a, b, x) <= 1e-12)


def test_bucket_coalescing_with_kept_factors(hip):
    # A burst of requests whose orders share one padded bucket (97..128 -> 128),
    # all keeping their factors: one worker takes the first, the rest queue up
    # behind it and are factorised as one padded stack. Each kept slice then
    # serves new right-hand sides, and every result equals the request alone.
    rng = np.random.default_rng(7)
    ns = [128, 97, 100, 110, 120, 101, 99, 127]
    mats = [cond_matrix(rng, n, 1e4) for n in ns]
    bs = [rng.standard_normal((n, 1 + i % 3)) for i, n in enumerate(ns)]
    srv = Server(budget_bytes=8e9, workers=1, max_batch=8, precision="fp32")
    ids = [srv.submit(a, b, key=100 + i, keep=True) for i, (a, b) in enumerate(zip(mats, bs))]
    res = [srv.wait(i) for i in ids]
    st = srv.stats()
    assert all(r["status"] == 0 for r in res)
    assert st["batched"] >= 2 and st["resident"] == len(ns), st
    for a, b, r in zip(mats, bs, res):
        assert np.all(oracle.backward_error(a, b, r["x"]) <= 1e-12)
        assert np.all(r["berr"] <= 1e-12)
    for i, a in enumerate(mats):
        b = rng.standard_normal((a.shape[0], 2))
        r = srv.wait(srv.submit(None, b, key=100 + i))
        assert r["status"] == 0
        assert np.all(oracle.backward_error(a, b, r["x"]) <= 1e-12)
    srv.close()
    for i in (1, 4, 7):
        s1 = Server(budget_bytes=8e9, workers=1, max_batch=1, precision="fp32")
        r = s1.wait(s1.submit(mats[i], bs[i]))
        s1.close()
        np.testing.assert_array_equal(r["x"], res[i]["x"])
