"""Writes the golden fixtures under tests/golden/ with NumPy/SciPy.

NumPP's oracle is live NumPy (tests/oracle.hpp:46-79 in the reference); these
fixtures freeze what NumPy 2.2 / SciPy 1.15 (LAPACK dgesv / dgetrf) return on
the same inputs, so the CPU oracle and the HIP path can be checked offline.

Fixtures (all .npz / .npy, loadable with allow_pickle=False):
  gesv_cases.npz      seeded systems of many sizes with pivoting forced:
                      A, B, X = np.linalg.solve(A, B), LAPACK ipiv (0-based
                      pivot row per step) and packed LU from
                      scipy.linalg.lu_factor.
  config1_x.npy       np.linalg.solve on the 256 x 64 x 64 stacked workload
                      (A = U[-1,1) + 64 I, b ~ N(0,1), default_rng(0)).
  matmul_cases.npz    seeded float64 / float32 GEMMs with numpy results.
  sgesv_cases.npz     seeded float32 systems with scipy.linalg.lu_factor /
                      lu_solve in float32 (LAPACK sgetrf / sgetrs): ipiv,
                      packed LU and X, for the float32 LapackVTable contract.
  tensor_cases.npz    np.linalg.tensorsolve / tensorinv on the reference's
                      tridiagonal 2x2x2x2 literal (test_tensor_solve.cpp:13-39)
                      and on a seeded 3x4x12 case.
  det_cases.npz       seeded general matrices (rows shuffled so pivoting
                      swaps) with np.linalg.det, np.linalg.slogdet and
                      np.linalg.inv (LAPACK dgetrf / dgesv).
  numpp_g15653376641833044941.npy
                      the reference's own frozen matmul golden
                      (arange(1,21).reshape(4,5) @ arange(2,17).reshape(5,3)),
                      copied as data from the reference's tests/golden/.

Run: python tests/golden/make_golden.py   (needs numpy + scipy, no GPU)
"""
from __future__ import annotations

import os
import shutil

import numpy as np
import scipy.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))


def gesv_cases():
    rng = np.random.default_rng(2024)
    sizes = [1, 2, 3, 4, 5, 7, 8, 13, 16, 17, 31, 32, 33, 48, 63, 64, 65, 100, 129]
    out = {}
    for idx, n in enumerate(sizes):
        nrhs = [1, 3, 8][idx % 3]
        # Random matrix whose rows are shuffled and scaled so that partial
        # pivoting has to swap rows; entries stay O(1) so the pivot gaps are wide.
        a = rng.uniform(-1.0, 1.0, (n, n))
        a += np.diag(rng.uniform(2.0, 3.0, n) * np.sign(rng.uniform(-1, 1, n)))
        perm = rng.permutation(n)
        a = a[perm]
        b = rng.standard_normal((n, nrhs))
        x = np.linalg.solve(a, b)
        lu, ipiv = sla.lu_factor(a)
        out[f"A_{n}"] = a
        out[f"B_{n}"] = b
        out[f"X_{n}"] = x
        out[f"LU_{n}"] = lu
        out[f"ipiv_{n}"] = ipiv.astype(np.int32)
    out["sizes"] = np.array(sizes, np.int32)
    np.savez(os.path.join(HERE, "gesv_cases.npz"), **out)


def config1():
    rng = np.random.default_rng(0)
    a = rng.uniform(-1.0, 1.0, (256, 64, 64)) + 64.0 * np.eye(64)
    b = rng.standard_normal((256, 64))
    x = np.linalg.solve(a, b[..., None])[..., 0]
    np.save(os.path.join(HERE, "config1_x.npy"), x)


def matmul_cases():
    rng = np.random.default_rng(1)
    out = {}
    shapes = [(1, 1, 1), (2, 3, 2), (37, 53, 29), (64, 64, 64), (65, 17, 130), (96, 80, 72)]
    for t, (m, k, n) in enumerate(shapes):
        a = rng.standard_normal((m, k))
        b = rng.standard_normal((k, n))
        out[f"A64_{t}"] = a
        out[f"B64_{t}"] = b
        out[f"C64_{t}"] = a @ b
        a32 = a.astype(np.float32)
        b32 = b.astype(np.float32)
        out[f"A32_{t}"] = a32
        out[f"B32_{t}"] = b32
        out[f"C32_{t}"] = a32 @ b32
    out["count"] = np.array([len(shapes)], np.int32)
    np.savez(os.path.join(HERE, "matmul_cases.npz"), **out)


def det_cases():
    rng = np.random.default_rng(2025)
    sizes = [1, 2, 3, 5, 8, 16, 33, 64, 65, 100, 129]
    out = {}
    for n in sizes:
        a = rng.uniform(-1.0, 1.0, (n, n)) + np.diag(rng.uniform(1.0, 2.0, n))
        a = a[rng.permutation(n)]
        out[f"A_{n}"] = a
        out[f"det_{n}"] = np.array(np.linalg.det(a))
        sign, logabs = np.linalg.slogdet(a)
        out[f"sign_{n}"] = np.array(sign)
        out[f"logabsdet_{n}"] = np.array(logabs)
        out[f"inv_{n}"] = np.linalg.inv(a)
    out["sizes"] = np.array(sizes, np.int32)
    np.savez(os.path.join(HERE, "det_cases.npz"), **out)


def sgesv_cases():
    rng = np.random.default_rng(2026)
    sizes = [3, 16, 40, 64, 65, 100, 200]
    out = {}
    for n in sizes:
        a = rng.uniform(-1.0, 1.0, (n, n)) + np.diag(rng.uniform(2.0, 3.0, n) * np.sign(rng.uniform(-1, 1, n)))
        a = a[rng.permutation(n)].astype(np.float32)
        b = rng.standard_normal((n, 3)).astype(np.float32)
        lu, ipiv = sla.lu_factor(a)
        assert lu.dtype == np.float32
        out[f"A_{n}"] = a
        out[f"B_{n}"] = b
        out[f"LU_{n}"] = lu
        out[f"ipiv_{n}"] = ipiv.astype(np.int32)
        out[f"X_{n}"] = sla.lu_solve((lu, ipiv), b)
    out["sizes"] = np.array(sizes, np.int32)
    np.savez(os.path.join(HERE, "sgesv_cases.npz"), **out)


def tensor_cases():
    m = np.array([4, 1, 0, 0, 1, 4, 1, 0, 0, 1, 4, 1, 0, 0, 1, 4.0]).reshape(2, 2, 2, 2)
    b = np.array([1, 2, 3, 4.0]).reshape(2, 2)
    rng = np.random.default_rng(2027)
    a2 = rng.standard_normal((3, 4, 12)) + np.eye(12).reshape(3, 4, 12) * 4.0
    b2 = rng.standard_normal((3, 4))
    np.savez(os.path.join(HERE, "tensor_cases.npz"), M=m, b=b, x=np.linalg.tensorsolve(m, b),
             Minv=np.linalg.tensorinv(m, 2), A2=a2, b2=b2, x2=np.linalg.tensorsolve(a2, b2),
             A2inv=np.linalg.tensorinv(a2, 2))


def reference_golden():
    src = "/root/reference/tests/golden/g15653376641833044941.npy"
    dst = os.path.join(HERE, "numpp_g15653376641833044941.npy")
    if os.path.exists(src):
        arr = np.load(src, allow_pickle=False)
        np.save(dst, arr)


if __name__ == "__main__":
    gesv_cases()
    config1()
    matmul_cases()
    det_cases()
    sgesv_cases()
    tensor_cases()
    reference_golden()
    print("goldens written to", HERE)
