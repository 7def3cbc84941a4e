"""Device buffers and streams through the C ABI (no torch in the product path)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class Stream:
    """A HIP stream owned by the library (default priority)."""

    def __init__(self):
        lib = _lib.require_device()
        h = ctypes.c_void_p()
        _lib.check(lib.npph_stream_create(ctypes.byref(h)))
        self.handle = h

    def sync(self):
        _lib.check(_lib.load().npph_stream_sync(self.handle))

    def close(self):
        if self.handle:
            _lib.load().npph_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A contiguous device array with numpy-style shape and dtype."""

    def __init__(self, shape, dtype=np.float64):
        lib = _lib.require_device()
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _lib.check(lib.npph_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p

    @classmethod
    def from_host(cls, arr, dtype=None):
        a = np.ascontiguousarray(arr, dtype=dtype or arr.dtype)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            _lib.check(_lib.load().npph_memcpy_htod(d.ptr, a.ctypes.data, d.nbytes, None))
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _lib.check(_lib.load().npph_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes, None))
        return out

    def copy_from(self, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        _lib.check(_lib.load().npph_memcpy_htod(self.ptr, a.ctypes.data, self.nbytes, None))

    def copy_from_device(self, other: "DeviceArray", stream=None):
        assert other.nbytes == self.nbytes
        _lib.check(_lib.load().npph_memcpy_dtod(self.ptr, other.ptr, self.nbytes, stream))

    def zero(self, stream=None):
        _lib.check(_lib.load().npph_memset(self.ptr, 0, self.nbytes, stream))

    @property
    def p(self):
        return self.ptr

    def free(self):
        if self.ptr:
            _lib.load().npph_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Timer:
    """HIP-event timer on a stream (None = the library stream)."""

    def __init__(self, stream=None):
        lib = _lib.require_device()
        self.stream = stream
        self.a, self.b = ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(lib.npph_event_create(ctypes.byref(self.a)))
        _lib.check(lib.npph_event_create(ctypes.byref(self.b)))

    def start(self):
        _lib.check(_lib.load().npph_event_record(self.a, self.stream))

    def stop(self) -> float:
        lib = _lib.load()
        _lib.check(lib.npph_event_record(self.b, self.stream))
        ms = ctypes.c_float()
        _lib.check(lib.npph_event_elapsed_ms(self.a, self.b, ctypes.byref(ms)))
        return ms.value


def device_sync():
    _lib.check(_lib.load().npph_device_sync())
