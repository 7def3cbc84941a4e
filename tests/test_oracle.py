"""Pins the CPU oracle (oracle/numpp_oracle.c) before anything is compared to it.

Each check is chosen so that a plausible mistake in the restatement fails it:
the reference's own literals and frozen golden, LAPACK/NumPy goldens written by
tests/golden/make_golden.py, exact rational elimination, and invariants.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle
from conftest import config1_inputs


# ---------------- matmul: backend.cpp:55-84 ----------------

def test_matmul_reference_literal():
    # tests/test_backend.cpp:46-55: [[1,2,3],[4,5,6]] @ [[1,2],[3,4],[5,6]]
    a = np.arange(1, 7, dtype=np.float64).reshape(2, 3)
    b = np.arange(1, 7, dtype=np.float64).reshape(3, 2)
    c = oracle.matmul(a, b)
    assert c[0, 0] == 22.0 and c[0, 1] == 28.0 and c[1, 0] == 49.0 and c[1, 1] == 64.0


def test_matmul_reference_frozen_golden(golden):
    # tests/test_backend.cpp:57-65 + tests/golden/g15653376641833044941.npy
    a = np.arange(1, 21, dtype=np.float64).reshape(4, 5)
    b = np.arange(2, 17, dtype=np.float64).reshape(5, 3)
    np.testing.assert_array_equal(oracle.matmul(a, b), golden("numpp_g15653376641833044941.npy"))


def test_matmul_numpy_goldens(golden):
    g = golden("matmul_cases.npz")
    for t in range(int(g["count"][0])):
        a, b, c = g[f"A64_{t}"], g[f"B64_{t}"], g[f"C64_{t}"]
        k = a.shape[1]
        bound = 2 * k * np.finfo(np.float64).eps * (np.abs(a) @ np.abs(b)) + 1e-300
        assert np.all(np.abs(oracle.matmul(a, b) - c) <= bound)
        a32, b32, c32 = g[f"A32_{t}"], g[f"B32_{t}"], g[f"C32_{t}"]
        bound32 = 2 * k * np.finfo(np.float32).eps * (np.abs(a32) @ np.abs(b32)) + 1e-30
        assert np.all(np.abs(oracle.matmul(a32, b32) - c32) <= bound32)


def test_matmul_exact_integers_and_order():
    # Small integers: every partial sum is exact, so the result must equal the
    # exact product; a transposed operand or dropped term would not.
    rng = np.random.default_rng(5)
    a = rng.integers(-8, 9, (13, 7)).astype(np.float64)
    b = rng.integers(-8, 9, (7, 11)).astype(np.float64)
    exact = np.array([[sum(int(a[i, p]) * int(b[p, j]) for p in range(7)) for j in range(11)] for i in range(13)])
    np.testing.assert_array_equal(oracle.matmul(a, b), exact.astype(np.float64))
    assert oracle.matmul(a, np.eye(7)).tolist() == a.tolist()


def test_dot2_exact():
    rng = np.random.default_rng(6)
    a = rng.integers(-1000, 1000, (4, 50)).astype(np.float64) * 2.0 ** 30
    b = rng.integers(-1000, 1000, (50, 3)).astype(np.float64) * 2.0 ** -30
    ref, mag = oracle.dot2_entries(a, b, [0, 1, 3], [0, 2, 1])
    for t, (i, j) in enumerate([(0, 0), (1, 2), (3, 1)]):
        exact = sum(Fraction(a[i, p]) * Fraction(b[p, j]) for p in range(50))
        assert ref[t] == float(exact)
        assert mag[t] == float(sum(abs(Fraction(a[i, p]) * Fraction(b[p, j])) for p in range(50)))


# ---------------- LU / solve: linalg.cpp:48-103 ----------------

def test_lu_pivots_match_lapack(golden):
    g = golden("gesv_cases.npz")
    for n in g["sizes"]:
        a = g[f"A_{n}"]
        lu, piv, steps, sgn, sing = oracle.lu_decompose(a)
        assert sing == -1
        # LAPACK dgetrf picks the first maximum (idamax), as NumPP's strict '>'.
        np.testing.assert_array_equal(steps, g[f"ipiv_{n}"])
        np.testing.assert_allclose(lu, g[f"LU_{n}"], rtol=0, atol=1e-12 * max(1.0, np.abs(g[f"LU_{n}"]).max()))
        # final permutation and sign are consistent with the swap sequence
        perm = np.arange(n)
        nswap = 0
        for k, p in enumerate(steps):
            if p != k:
                perm[[k, p]] = perm[[p, k]]
                nswap += 1
        np.testing.assert_array_equal(piv, perm)
        assert sgn == (-1.0) ** nswap
        L = np.tril(lu, -1) + np.eye(n)
        U = np.triu(lu)
        np.testing.assert_allclose(L @ U, a[perm], rtol=0, atol=1e-13 * n * np.abs(a).max())


def test_solve_numpy_goldens(golden):
    g = golden("gesv_cases.npz")
    for n in g["sizes"]:
        a, b, x = g[f"A_{n}"], g[f"B_{n}"], g[f"X_{n}"]
        xo = oracle.solve(a, b)
        np.testing.assert_allclose(xo, x, rtol=1e-9, atol=1e-11)  # tests/test_batched_linalg.cpp:41,50


def test_config1_golden(golden):
    a, b = config1_inputs()
    x, info = oracle.batched_solve(a, b)
    assert np.all(info == -1)
    xg = golden("config1_x.npy")
    assert np.abs(x - xg).max() <= 1e-13 * np.abs(xg).max()


def _exact_gepp(a, b):
    """Partial-pivoting elimination in exact rationals, NumPP's pivot rule."""
    n = len(a)
    A = [[Fraction(v) for v in row] for row in a]
    B = [[Fraction(v) for v in row] for row in b]
    steps = []
    for k in range(n):
        pr, best = k, abs(A[k][k])
        for i in range(k + 1, n):
            if abs(A[i][k]) > best:
                best, pr = abs(A[i][k]), i
        steps.append(pr)
        if best == 0:
            return None, steps
        A[k], A[pr] = A[pr], A[k]
        B[k], B[pr] = B[pr], B[k]
        for i in range(k + 1, n):
            f = A[i][k] / A[k][k]
            for j in range(k, n):
                A[i][j] -= f * A[k][j]
            for j in range(len(B[0])):
                B[i][j] -= f * B[k][j]
    X = [[Fraction(0)] * len(B[0]) for _ in range(n)]
    for i in range(n - 1, -1, -1):
        for j in range(len(B[0])):
            s = B[i][j] - sum(A[i][q] * X[q][j] for q in range(i + 1, n))
            X[i][j] = s / A[i][i]
    return X, steps


def test_solve_exact_small():
    rng = np.random.default_rng(7)
    for n in (1, 2, 3, 4):
        for _ in range(20):
            a = rng.integers(-9, 10, (n, n)).astype(np.float64)
            b = rng.integers(-9, 10, (n, 2)).astype(np.float64)
            X, steps = _exact_gepp(a, b)
            lu, piv, st, sgn, sing = oracle.lu_decompose(a)
            if X is None:
                assert sing >= 0
                continue
            assert sing == -1
            assert st.tolist() == steps
            xo = oracle.solve(a, b)
            xe = np.array([[float(v) for v in row] for row in X])
            assert np.all(np.abs(xo - xe) <= 1e-12 * np.maximum(1.0, np.abs(xe)))


def test_special_cases():
    # identity, permutation, triangular
    for n in (1, 5, 17):
        b = np.arange(1.0, n + 1.0)
        np.testing.assert_array_equal(oracle.solve(np.eye(n), b), b)
        p = np.eye(n)[::-1]
        np.testing.assert_array_equal(oracle.solve(p, b), b[::-1])
        t = np.triu(np.ones((n, n))) + np.eye(n)
        np.testing.assert_allclose(t @ oracle.solve(t, b), b, rtol=1e-13)
    # tie: equal magnitudes -> the first row wins (strict '>', linalg.cpp:55-56)
    a = np.array([[1.0, 2.0], [-1.0, 3.0]])
    _, _, steps, _, _ = oracle.lu_decompose(a)
    assert steps[0] == 0
    a = np.array([[0.5, 2.0], [-1.0, 3.0], ])
    _, _, steps, _, _ = oracle.lu_decompose(a)
    assert steps[0] == 1


def test_singular_exact_zero_rule():
    a = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [1.0, 0.0, 1.0]])
    # column 0 fine; after elimination column 1 has (0, -2): pivot -2 -> fine;
    # rank 2 means a zero pivot appears at step 2
    _, _, steps, _, sing = oracle.lu_decompose(a)
    assert sing == 2
    z = np.zeros((4, 4))
    assert oracle.lu_decompose(z)[4] == 0
    with pytest.raises(oracle.SingularMatrix):
        oracle.solve(z, np.ones(4))
    a = np.eye(5)
    a[3, 3] = 0.0
    assert oracle.lu_decompose(a)[4] == 3


def test_batched_semantics():
    # tests/test_batched_linalg.cpp:35-51: stack3 = arange(12).reshape(3,2,2)
    A = np.arange(12.0).reshape(3, 2, 2)
    b = np.arange(1.0, 7.0).reshape(3, 2)
    x, info = oracle.batched_solve(A, b)
    assert x.shape == (3, 2) and np.all(info == -1)
    np.testing.assert_allclose(x, np.linalg.solve(A, b[..., None])[..., 0], rtol=1e-9, atol=1e-11)
    B = np.arange(1.0, 13.0).reshape(3, 2, 2)
    x2, _ = oracle.batched_solve(A, B)
    np.testing.assert_allclose(x2, np.linalg.solve(A, B), rtol=1e-9, atol=1e-11)


def test_backward_error():
    rng = np.random.default_rng(8)
    a = rng.standard_normal((30, 30))
    b = rng.standard_normal((30, 2))
    x = oracle.solve(a, b)
    eta = oracle.backward_error(a, b, x)
    assert np.all(eta < 1e-14)
    # perturbing x by delta raises eta to about |A delta| / (|A||x| + |b|)
    xp = x.copy()
    xp[3, 1] += 1e-6
    eta2 = oracle.backward_error(a, b, xp)
    r = b[:, 1] - a @ xp[:, 1]
    expect = np.abs(r).max() / (np.abs(a).sum(1).max() * np.abs(xp[:, 1]).max() + np.abs(b[:, 1]).max())
    assert eta2[0] < 1e-14 and abs(eta2[1] - expect) <= 1e-6 * expect


def test_planted_mistake_is_caught():
    """A restatement with the wrong tie rule ('>=' picks the last maximum) must
    fail the LAPACK pivot comparison on a tie: the check has teeth."""
    a = np.array([[1.0, 1.0, 0.0], [1.0, 2.0, 1.0], [-1.0, 0.0, 3.0]])
    _, _, steps, _, _ = oracle.lu_decompose(a)
    import scipy.linalg as sla
    assert steps.tolist() == sla.lu_factor(a)[1].tolist()
    # the mutated rule would choose row 2 at step 0
    col = np.abs(a[:, 0])
    last_max = int(np.flatnonzero(col == col.max())[-1])
    assert last_max != steps[0]


# ---------------- det / slogdet / inv: linalg.cpp:105-115, 393-439 ----------------

def _exact_det_inv(a):
    """Determinant and inverse in exact rationals (Gauss-Jordan)."""
    n = len(a)
    A = [[Fraction(v) for v in row] + [Fraction(int(i == j)) for j in range(n)] for i, row in enumerate(a)]
    det = Fraction(1)
    for k in range(n):
        p = next((i for i in range(k, n) if A[i][k] != 0), None)
        if p is None:
            return Fraction(0), None
        if p != k:
            A[k], A[p] = A[p], A[k]
            det = -det
        det *= A[k][k]
        piv = A[k][k]
        A[k] = [v / piv for v in A[k]]
        for i in range(n):
            if i != k and A[i][k] != 0:
                f = A[i][k]
                A[i] = [vi - f * vk for vi, vk in zip(A[i], A[k])]
    return det, [row[n:] for row in A]


def test_det_inv_numpy_goldens(golden):
    g = golden("det_cases.npz")
    for n in g["sizes"]:
        a = g[f"A_{n}"]
        d = oracle.det(a)
        assert abs(d - g[f"det_{n}"]) <= 1e-12 * n * abs(g[f"det_{n}"])
        s, la = oracle.slogdet(a)
        assert s == g[f"sign_{n}"] and abs(la - g[f"logabsdet_{n}"]) <= 1e-12 * n * max(1.0, abs(la))
        assert abs(d - s * np.exp(la)) <= 1e-12 * n * abs(d)
        ai = oracle.inv(a)
        np.testing.assert_allclose(ai, g[f"inv_{n}"], rtol=0, atol=1e-11 * np.abs(g[f"inv_{n}"]).max())
        np.testing.assert_allclose(a @ ai, np.eye(n), rtol=0, atol=1e-12 * n)


def test_det_inv_exact_small():
    rng = np.random.default_rng(11)
    for n in (1, 2, 3, 4):
        for _ in range(25):
            a = rng.integers(-9, 10, (n, n)).astype(np.float64)
            de, ie = _exact_det_inv(a.tolist())
            d = oracle.det(a)
            s, la = oracle.slogdet(a)
            if de == 0:
                assert d == 0.0 and s == 0.0 and la == -np.inf
                with pytest.raises(oracle.SingularMatrix):
                    oracle.inv(a)
                continue
            assert abs(d - float(de)) <= 1e-13 * abs(float(de))
            assert s == (1.0 if de > 0 else -1.0)
            assert abs(la - np.log(abs(float(de)))) <= 1e-13 * max(1.0, abs(la))
            xe = np.array([[float(v) for v in row] for row in ie])
            assert np.all(np.abs(oracle.inv(a) - xe) <= 1e-13 * np.maximum(1.0, np.abs(xe)))


def test_det_special_cases():
    for n in (0, 1, 6, 40):
        assert oracle.det(np.eye(n)) == 1.0 and oracle.slogdet(np.eye(n)) == (1.0, 0.0)
    # permutation matrices: the sign of the permutation, exactly
    rng = np.random.default_rng(12)
    for n in (2, 5, 9):
        p = rng.permutation(n)
        inversions = sum(1 for i in range(n) for j in range(i + 1, n) if p[i] > p[j])
        assert oracle.det(np.eye(n)[p]) == (-1.0) ** inversions
    # upper triangular: no row swaps, the diagonal product in row order
    t = np.triu(rng.uniform(-1, 1, (7, 7))) + np.diag(rng.uniform(1, 2, 7))
    want = 1.0
    for v in np.diag(t):
        want *= v
    assert oracle.det(t) == want
    s, la = oracle.slogdet(t)
    assert s == np.sign(want) and la == sum(np.log(np.abs(np.diag(t))).tolist())
    # exact-zero singularity (linalg.cpp:57) -> 0, (0, -inf), inv raises
    a = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [1.0, 0.0, 1.0]])
    assert oracle.det(a) == 0.0 and oracle.slogdet(a) == (0.0, -np.inf)
    with pytest.raises(oracle.SingularMatrix):
        oracle.inv(a)


def test_numpp_equivalent_timing_paths_compute_the_same():
    # the "NumPP-equivalent" CPU baselines add NumPP's ndarray copies
    # (ndarray.cpp:216-231, 257-285) around the same kernels: same values
    a, b = config1_inputs()
    x1, _ = oracle.batched_solve(a[:8], b[:8])
    np.testing.assert_array_equal(oracle.batched_solve_numpp(a[:8], b[:8]), x1)
    rng = np.random.default_rng(13)
    A, B = rng.standard_normal((30, 20)), rng.standard_normal((20, 10))
    np.testing.assert_array_equal(oracle.matmul_numpp(A, B), oracle.matmul(A, B))


# ---------------- element-wise ufuncs and reductions: ufunc.cpp:450-472 ----------------

def _specials(dt):
    f = np.finfo(dt)
    return np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, f.tiny, -f.tiny, f.smallest_subnormal,
                     f.max, -f.max, 3.0, 1e-3, 7.5], dtype=dt)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_elementwise_matches_numpy_ieee(dt):
    rng = np.random.default_rng(21)
    s = _specials(dt)
    a = np.concatenate([np.repeat(s, s.size), rng.standard_normal(500).astype(dt)])
    b = np.concatenate([np.tile(s, s.size), rng.standard_normal(500).astype(dt)])
    with np.errstate(all="ignore"):
        want = {"add": a + b, "subtract": a - b, "multiply": a * b, "divide": a / b, "negative": -a,
                "sqrt": np.sqrt(a)}
    for name, w in want.items():
        got = oracle.elementwise(name, a, None if name in ("negative", "sqrt") else b)
        assert got.dtype == dt
        np.testing.assert_array_equal(np.isnan(got), np.isnan(w))
        ok = ~np.isnan(w)
        assert np.array_equal(got[ok].view(np.uint32 if dt == np.float32 else np.uint64),
                              w[ok].view(np.uint32 if dt == np.float32 else np.uint64)), name


def test_reduce_sequential_order():
    # exact on small integers; the sequential order is visible where rounding
    # depends on it: ((1e16 + 1) + 1) - 1e16 = 0 in order, 2 exactly
    x = np.arange(1.0, 1001.0)
    assert oracle.reduce("sum", x) == 500500.0
    assert oracle.reduce("sum", np.array([1e16, 1.0, 1.0, -1e16])) == 0.0
    assert oracle.reduce("prod", np.array([2.0, 0.5, 4.0, 3.0])) == 12.0
    assert oracle.reduce("sum", np.array([], np.float64)) == 0.0 and oracle.reduce("prod", np.array([])) == 1.0
    f = np.full(2 ** 25, 1.0, np.float32)
    # float32 accumulation stalls at 2^24 (1 + 2^24 rounds to 2^24): the dtype is the accumulator's
    assert oracle.reduce("sum", f) == np.float32(2 ** 24)
