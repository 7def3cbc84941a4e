"""The benchmark workloads at their real sizes, checked against the oracle.

Mixed-precision stack (n = 2048, cond 1e4, 8 right-hand sides): four systems
A = U diag(s) V^T built like the benchmark's (Haar U, V; s = logspace(0, -4)),
solved in FP32 + refinement, split-BF16 (3 and 6 products) + refinement and
plain FP64. Bounds: the device-reported backward error and the oracle's
Dot2-residual backward error of every column are <= 1e-12; the forward error
against the oracle's FP64 LU solution (two systems) is <= 1e-12 * cond.

Uneven batch (the benchmark's full plan): 20,000 systems, n ~ U{8..256},
nrhs ~ U{1..16}, A = U[-1,1) + n I, B ~ N(0,1), PCG64 seed 4, in FP64 and FP32
+ refinement. Bounds: every reported backward error and every host-recomputed
one (numpy residual, rounding below 256 u |A||x| ~ 3e-14) is <= 1e-12; a
stratified sample of 96 systems against oracle.solve within 1e-12 relative
(these systems have cond_inf < 4) and the oracle's Dot2 backward error; 24
systems solved alone (their own plan) agree with the full plan within 1e-13.

Serving trace: requests with n from 16 to 4096 and cond up to 1e8 (FP32
refinement cannot reach 1e-12 there, so the FP64 fallback must fire), reuse of
resident factors, the 32e9-byte budget. Bounds: every result's Dot2 backward
error <= 1e-12; results bitwise the same alone and in the mix; the budget's
peak is respected.
"""
import numpy as np
import pytest

import oracle
from numpp_amd import linalg as nla
from numpp_amd.device import DeviceArray
from numpp_amd.serve import Server
from numpp_amd.vbatch import VPlan, solve_uneven

pytestmark = pytest.mark.gpu


def haar(rng, n):
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))


# ---------------------------------------------------------------- n = 2048

@pytest.fixture(scope="module")
def mixed_stack():
    rng = np.random.default_rng(3)
    n, nrhs = 2048, 8
    s = np.logspace(0.0, -4.0, n)
    us = [haar(rng, n) * s for _ in range(2)]
    vt = [np.ascontiguousarray(haar(rng, n).T) for _ in range(2)]
    a = np.stack([us[i] @ vt[j] for i in range(2) for j in range(2)])
    b = rng.standard_normal((4, n, nrhs))
    xo = [oracle.solve(a[i], b[i]) for i in (0, 3)]
    return a, b, xo


@pytest.mark.parametrize("mode", ["fp32", "bf16x6", "bf16x3", "fp64"])
def test_mixed_2048(hip, mixed_stack, mode):
    a, b, xo = mixed_stack
    prec, terms = ("bf16", int(mode[5:])) if mode.startswith("bf16") else (mode, 6)
    x, berr, iters, info = nla.solve_refined(a, b, precision=prec, bf16_terms=terms)
    print(f"{mode}: iters {iters.tolist()} berr_max {berr.max():.2e}", flush=True)
    assert np.all(info == 0)
    assert np.all(berr <= 1e-12), berr.max()
    for i in range(a.shape[0]):
        eta = oracle.backward_error(a[i], b[i], x[i])
        assert np.all(eta <= 1e-12), (i, eta.max())
    for i, ref in zip((0, 3), xo):
        fe = np.abs(x[i] - ref).max() / np.abs(ref).max()
        assert fe <= 1e-12 * 1e4, (i, fe)
    if mode == "fp32":
        # one system alone against the same system inside the stack
        x1, berr1, _, _ = nla.solve_refined(a[2:3], b[2:3], precision=prec, bf16_terms=terms)
        d = np.abs(x1[0] - x[2]).max() / np.abs(x[2]).max()
        print(f"alone vs mix: max rel diff {d:.3e} (bitwise: {np.array_equal(x1[0], x[2])})", flush=True)
        assert d <= 1e-12 * 1e4


# ---------------------------------------------------------------- uneven batch

@pytest.fixture(scope="module")
def uneven():
    S = 20000
    rng = np.random.default_rng(4)
    ns = rng.integers(8, 257, size=S).astype(np.int32)
    nr = rng.integers(1, 17, size=S).astype(np.int32)
    offA = np.concatenate([[0], np.cumsum(ns.astype(np.int64) ** 2)])
    offB = np.concatenate([[0], np.cumsum(ns.astype(np.int64) * nr)])
    A = np.empty(int(offA[-1]))
    B = np.empty(int(offB[-1]))
    for i in range(S):
        n = int(ns[i])
        a = rng.uniform(-1.0, 1.0, (n, n))
        a[np.diag_indices(n)] += n
        A[offA[i]:offA[i + 1]] = a.ravel()
        B[offB[i]:offB[i + 1]] = rng.standard_normal((n, int(nr[i]))).ravel()
    return ns, nr, offA, offB, A, B


def _sys(U, i):
    ns, nr, offA, offB, A, B = U
    n, r = int(ns[i]), int(nr[i])
    return A[offA[i]:offA[i + 1]].reshape(n, n), B[offB[i]:offB[i + 1]].reshape(n, r)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_uneven_20000(hip, uneven, precision):
    ns, nr, offA, offB, A, B = uneven
    S = len(ns)
    plan = VPlan(ns, nr, precision=precision, tol=1e-12, max_iter=10, fallback=True, autotune=True)
    dA, dB = DeviceArray.from_host(A), DeviceArray.from_host(B)
    dX = DeviceArray(B.shape)
    dE = DeviceArray((int(nr.sum()),))
    dI = DeviceArray((S,), np.int32)
    dF = DeviceArray((S,), np.int32)
    plan.execute(dA, dB, dX, dE, dI, dF, use_graph=False)
    X, E, it, info = dX.to_host(), dE.to_host(), dI.to_host(), dF.to_host()
    plan.close()
    print(f"{precision}: berr_max {E.max():.2e}, iters {it.min()}..{it.max()}", flush=True)
    assert np.all(info == 0)
    assert np.all(E <= 1e-12), E.max()
    assert np.all(it == 0) if precision == "fp64" else np.all(it != 0)
    worst = 0.0
    for i in range(S):
        a, b = _sys(uneven, i)
        x = X[offB[i]:offB[i + 1]].reshape(b.shape)
        r = b - a @ x
        eta = np.abs(r).max(axis=0) / (np.abs(a).sum(axis=1).max() * np.abs(x).max(axis=0) + np.abs(b).max(axis=0))
        worst = max(worst, float(eta.max()))
    assert worst <= 1e-12, worst
    order = np.argsort(ns, kind="stable")
    sample = order[np.linspace(0, S - 1, 96).astype(int)]
    for i in sample:
        a, b = _sys(uneven, i)
        x = X[offB[i]:offB[i + 1]].reshape(b.shape)
        xo = oracle.solve(a, b)
        assert np.abs(x - xo).max() <= 1e-12 * np.abs(xo).max(), i
        assert np.all(oracle.backward_error(a, b, x) <= 1e-12), i
    # alone (a plan of 24 systems) against the full plan
    few = sample[::4]
    xs, _, _, _ = solve_uneven([_sys(uneven, i)[0] for i in few], [_sys(uneven, i)[1] for i in few],
                               precision=precision)
    for i, xa in zip(few, xs):
        x = X[offB[i]:offB[i + 1]].reshape(xa.shape)
        assert np.abs(xa - x).max() <= 1e-13 * np.abs(x).max(), i


# ---------------------------------------------------------------- serving trace

@pytest.fixture(scope="module")
def serve_trace():
    rng = np.random.default_rng(5)
    spec = [(4096, 1e8, 3), (2500, 1e6, 16), (1024, 1e7, 1), (700, 1e3, 64), (2048, 1e5, 8), (300, 1e2, 5),
            (129, 1e4, 2), (64, 1e8, 1), (33, 1e3, 7), (16, 1e2, 1), (300, 1e4, 5), (129, 1e3, 9)]
    mats, rhss, keys, keeps = [], [], [], []
    for i, (n, cond, r) in enumerate(spec):
        s = np.logspace(0.0, -np.log10(cond), n)
        mats.append((haar(rng, n) * s) @ haar(rng, n).T)
        rhss.append(rng.standard_normal((n, r)))
        keys.append(i + 1)
        keeps.append(1)
    # reuse requests: new right-hand sides against resident factors
    for j in (0, 1, 4, 6):
        mats.append(None)
        rhss.append(rng.standard_normal((spec[j][0], 4)))
        keys.append(j + 1)
        keeps.append(0)
    gaps = rng.exponential(1.0 / 50.0, len(mats))
    t = np.cumsum(gaps) - gaps[0]
    return mats, rhss, keys, keeps, t


def test_serving_trace(hip, serve_trace):
    mats, rhss, keys, keeps, t = serve_trace
    srv = Server(budget_bytes=32e9, workers=4, max_batch=8, precision="fp32")
    out = srv.replay(t, mats, rhss, keys, keeps)
    st = srv.stats()
    srv.close()
    assert out["status"] == 0 and np.all(out["req_status"] == 0), out["req_status"]
    assert st["peak_bytes"] <= 32e9
    owner = {k: m for k, m in zip(keys, mats) if m is not None}
    for i, (m, b) in enumerate(zip(mats, rhss)):
        a = owner[keys[i]] if m is None else m
        eta = oracle.backward_error(a, b, out["x"][i])
        assert np.all(eta <= 1e-12), (i, a.shape[0], eta.max())
        assert np.all(out["berr"][i] <= 1e-12), (i, out["berr"][i])
    # kappa 1e8 at n = 4096 is beyond FP32 refinement: it was re-solved in FP64
    assert out["iters"][0] == -1
    print(f"iters {out['iters'].tolist()}, p99 {np.percentile(out['latency_ms'], 99):.1f} ms", flush=True)
    # each one-shot result alone equals the in-mix result bitwise
    for i in (0, 2, 5, 9):
        s1 = Server(budget_bytes=32e9, workers=1, max_batch=1, precision="fp32")
        r = s1.wait(s1.submit(mats[i], rhss[i]))
        s1.close()
        assert r["status"] == 0
        np.testing.assert_array_equal(r["x"], out["x"][i])
