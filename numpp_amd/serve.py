"""Serving loop (npph_server_* in include/numpp_hip.h): requests of different
n and nrhs served as they arrive, with byte-budget admission, same-order
coalescing and resident factorisations for repeat right-hand sides."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .linalg import PRECISIONS


class Server:
    def __init__(self, budget_bytes=32e9, workers=4, max_batch=8, precision="fp32", tol=1e-12, max_iter=10,
                 fallback=True, bf16_terms=6):
        lib = _lib.require_device()
        self.cfg = _lib.ServeConfig(float(budget_bytes), int(workers), int(max_batch),
                                    PRECISIONS[precision] if isinstance(precision, str) else int(precision),
                                    int(max_iter), float(tol), 1 if fallback else 0, int(bf16_terms))
        h = ctypes.c_void_p()
        _lib.check(lib.npph_server_create(ctypes.byref(self.cfg), ctypes.byref(h)))
        self.handle = h
        self._keep = {}  # request id -> host buffers kept alive until wait

    def submit(self, A, B, key=0, keep=False):
        """A: (n, n) float64 or None to reuse resident factors under `key`; B: (n,) or (n, nrhs)."""
        B2 = np.ascontiguousarray(B, np.float64)
        vec = B2.ndim == 1
        if vec:
            B2 = B2[:, None].copy()
        n, nrhs = B2.shape
        A2 = None if A is None else np.ascontiguousarray(A, np.float64)
        X = np.empty_like(B2)
        berr = np.empty(nrhs)
        rid = ctypes.c_int64()
        _lib.check(_lib.load().npph_server_submit(self.handle, n, nrhs, None if A2 is None else A2.ctypes.data,
                                                  B2.ctypes.data, X.ctypes.data, berr.ctypes.data, int(key),
                                                  1 if keep else 0, ctypes.byref(rid)))
        self._keep[rid.value] = (A2, B2, X, berr, vec)
        return rid.value

    def wait(self, rid):
        """Returns dict(x, berr, iters, info, latency_ms, queue_ms, status)."""
        it, inf = ctypes.c_int32(), ctypes.c_int32()
        lat, qms = ctypes.c_double(), ctypes.c_double()
        st = _lib.load().npph_server_wait(self.handle, rid, ctypes.byref(it), ctypes.byref(inf), ctypes.byref(lat),
                                          ctypes.byref(qms))
        A2, B2, X, berr, vec = self._keep.pop(rid)
        return {"x": X[:, 0] if vec else X, "berr": berr, "iters": it.value, "info": inf.value,
                "latency_ms": lat.value, "queue_ms": qms.value, "status": st}

    def stats(self):
        vals = [ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int64(),
                ctypes.c_int64(), ctypes.c_int64()]
        _lib.check(_lib.load().npph_server_stats(self.handle, *[ctypes.byref(v) for v in vals]))
        keys = ["completed", "resident", "bytes_in_use", "peak_bytes", "batched", "fallbacks", "evictions"]
        return {k: v.value for k, v in zip(keys, vals)}

    def replay(self, t_arrive, mats, rhss, keys, keeps):
        """Submit request i at t_arrive[i] seconds (C++ submitter thread).
        mats[i] may be None (reuse). Returns per-request arrays and wall time."""
        cnt = len(t_arrive)
        t = np.ascontiguousarray(t_arrive, np.float64)
        ns = np.array([r.shape[0] for r in rhss], np.int32)
        nr = np.array([1 if r.ndim == 1 else r.shape[1] for r in rhss], np.int32)
        Bs = [np.ascontiguousarray(r.reshape(r.shape[0], -1), np.float64) for r in rhss]
        Xs = [np.empty_like(b) for b in Bs]
        Es = [np.empty(b.shape[1]) for b in Bs]
        As = [None if a is None else np.ascontiguousarray(a, np.float64) for a in mats]
        PA = (ctypes.c_void_p * cnt)(*[None if a is None else a.ctypes.data for a in As])
        PB = (ctypes.c_void_p * cnt)(*[b.ctypes.data for b in Bs])
        PX = (ctypes.c_void_p * cnt)(*[x.ctypes.data for x in Xs])
        PE = (ctypes.c_void_p * cnt)(*[e.ctypes.data for e in Es])
        K = np.ascontiguousarray(keys, np.uint64)
        KP = np.ascontiguousarray(keeps, np.int32)
        lat = np.zeros(cnt)
        its = np.zeros(cnt, np.int32)
        inf = np.zeros(cnt, np.int32)
        sts = np.zeros(cnt, np.int32)
        wall = ctypes.c_double()
        st = _lib.load().npph_server_replay(self.handle, cnt, t.ctypes.data, ns.ctypes.data, nr.ctypes.data,
                                            ctypes.cast(PA, ctypes.c_void_p), ctypes.cast(PB, ctypes.c_void_p),
                                            ctypes.cast(PX, ctypes.c_void_p), ctypes.cast(PE, ctypes.c_void_p),
                                            K.ctypes.data, KP.ctypes.data, lat.ctypes.data, its.ctypes.data,
                                            inf.ctypes.data, sts.ctypes.data, ctypes.byref(wall))
        return {"status": st, "latency_ms": lat, "iters": its, "info": inf, "req_status": sts, "wall_s": wall.value,
                "x": Xs, "berr": Es}

    def close(self):
        if self.handle:
            _lib.load().npph_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
