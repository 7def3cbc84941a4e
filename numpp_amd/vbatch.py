"""Uneven batches: systems of different order and right-hand-side count solved
in one planned call (npph_vbatch_* in include/numpp_hip.h)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray
from .linalg import refine_opts


def pack(mats, rhss):
    """Packed host buffers (A, B) and (n, nrhs) arrays for a list of systems."""
    n = np.array([m.shape[0] for m in mats], np.int32)
    nr = np.array([1 if r.ndim == 1 else r.shape[1] for r in rhss], np.int32)
    A = np.concatenate([np.ascontiguousarray(m, np.float64).ravel() for m in mats])
    B = np.concatenate([np.ascontiguousarray(r, np.float64).ravel() for r in rhss])
    return A, B, n, nr


def unpack(X, n, nr):
    out, off = [], 0
    for ni, ri in zip(n, nr):
        out.append(X[off:off + ni * ri].reshape(ni, ri))
        off += ni * ri
    return out


class VPlan:
    """A bucketed plan for fixed (n_i, nrhs_i); execute() may be called many times."""

    def __init__(self, n, nrhs, precision="fp64", tol=1e-12, max_iter=10, fallback=True, bf16_terms=6,
                 autotune=True):
        lib = _lib.require_device()
        self.n = np.ascontiguousarray(n, np.int32)
        self.nrhs = np.ascontiguousarray(nrhs, np.int32)
        self.count = len(self.n)
        self.opts = refine_opts(precision, tol, max_iter, fallback, bf16_terms)
        h = ctypes.c_void_p()
        _lib.check(lib.npph_vbatch_plan_create(self.count, self.n.ctypes.data, self.nrhs.ctypes.data,
                                               ctypes.byref(self.opts), 1 if autotune else 0, ctypes.byref(h)))
        self.handle = h

    def buckets(self):
        lib = _lib.load()
        nb = ctypes.c_int()
        cap = 64
        order = np.zeros(cap, np.int32)
        cnt = np.zeros(cap, np.int64)
        rp = np.zeros(cap, np.int32)
        _lib.check(lib.npph_vbatch_plan_info(self.handle, ctypes.byref(nb), order.ctypes.data, cnt.ctypes.data,
                                             rp.ctypes.data, cap))
        k = nb.value
        return list(zip(order[:k].tolist(), cnt[:k].tolist(), rp[:k].tolist()))

    def execute(self, dA: DeviceArray, dB: DeviceArray, dX: DeviceArray, dberr: DeviceArray | None = None,
                diters: DeviceArray | None = None, dinfo: DeviceArray | None = None, use_graph=True, stream=None):
        lib = _lib.load()
        _lib.check(lib.npph_vbatch_execute(self.handle, dA.ptr, dB.ptr, dX.ptr, dberr.ptr if dberr else None,
                                           diters.ptr if diters else None, dinfo.ptr if dinfo else None,
                                           1 if use_graph else 0, stream))

    def close(self):
        if self.handle:
            _lib.load().npph_vbatch_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_uneven(mats, rhss, precision="fp64", use_graph=False, **kw):
    """Host convenience: returns (list of X_i, berr (sum nrhs), iters, info)."""
    A, B, n, nr = pack(mats, rhss)
    plan = VPlan(n, nr, precision, **kw)
    dA, dB = DeviceArray.from_host(A), DeviceArray.from_host(B)
    dX = DeviceArray(B.shape)
    dE = DeviceArray((int(nr.sum()),))
    dI = DeviceArray((len(n),), np.int32)
    dF = DeviceArray((len(n),), np.int32)
    plan.execute(dA, dB, dX, dE, dI, dF, use_graph=use_graph)
    X = dX.to_host()
    res = (unpack(X, n, nr), dE.to_host(), dI.to_host(), dF.to_host())
    plan.close()
    return res
