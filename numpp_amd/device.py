"""Device buffers and streams through the C ABI (no torch in the product path),
and their zero-copy exchange through DLPack on ROCm devices.

DLPack: NumPP's vendored dlpack.h knows only kDLCPU / kDLCUDA / kDLCUDAHost /
kDLOpenCL (include/numpp/interop/dlpack.h:9) and its to_dlpack hard-codes
kDLCPU (src/interop/interop.cpp:35). DeviceArray speaks the DLPack protocol
(__dlpack__ / __dlpack_device__, from_dlpack) with device type kDLROCM = 10,
the value the DLPack specification assigns to AMD GPUs, so device-resident
arrays move between this backend and ROCm frameworks (torch on ROCm reports
kDLROCM) without host copies.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

kDLCPU, kDLROCM = 1, 10
_DL_CODES = {"i": 0, "u": 1, "f": 2, "c": 5, "b": 6}


class DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int32), ("device_id", ctypes.c_int32)]


class DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class DLTensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("device", DLDevice), ("ndim", ctypes.c_int32),
                ("dtype", DLDataType), ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


class DLManagedTensor(ctypes.Structure):
    pass


_DELETER = ctypes.CFUNCTYPE(None, ctypes.POINTER(DLManagedTensor))
DLManagedTensor._fields_ = [("dl_tensor", DLTensor), ("manager_ctx", ctypes.c_void_p), ("deleter", _DELETER)]

_capi = ctypes.pythonapi
_capi.PyCapsule_New.restype = ctypes.py_object
_capi.PyCapsule_New.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
_capi.PyCapsule_GetPointer.restype = ctypes.c_void_p
_capi.PyCapsule_GetPointer.argtypes = [ctypes.py_object, ctypes.c_char_p]
_capi.PyCapsule_SetName.restype = ctypes.c_int
_capi.PyCapsule_SetName.argtypes = [ctypes.py_object, ctypes.c_char_p]
_capi.PyCapsule_IsValid.restype = ctypes.c_int
_capi.PyCapsule_IsValid.argtypes = [ctypes.py_object, ctypes.c_char_p]

# Exported tensors alive until their consumer calls the deleter: id -> (struct,
# shape array, owning array). The deleter drops the entry.
_exported = {}


@_DELETER
def _dl_deleter(mt):
    _exported.pop(ctypes.addressof(mt.contents), None)


_CAPSULE_DTOR = ctypes.CFUNCTYPE(None, ctypes.c_void_p)


# the capsule C API on a raw PyObject* (the destructor gets the capsule
# being destroyed, which must not be wrapped in a new Python reference)
_IsValidRaw = ctypes.PYFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p)(("PyCapsule_IsValid", _capi))
_GetPointerRaw = ctypes.PYFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p)(("PyCapsule_GetPointer", _capi))


@_CAPSULE_DTOR
def _capsule_dtor(cap):
    # an unconsumed capsule (still named "dltensor") releases its tensor
    if _IsValidRaw(cap, b"dltensor"):
        _exported.pop(_GetPointerRaw(cap, b"dltensor"), None)


def current_device() -> int:
    """The device the library runs on (npph_device_id)."""
    return int(_lib.load().npph_device_id())


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

    def __init__(self, shape, dtype=np.float64, _ptr=None, _owner=None):
        lib = _lib.require_device()
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self.device = current_device()
        self._owner = _owner  # a view of memory owned elsewhere (DLPack import)
        if _ptr is not None:
            self.ptr = ctypes.c_void_p(_ptr)
            return
        p = ctypes.c_void_p()
        _lib.check(lib.npph_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p

    # ---- DLPack (kDLROCM) ----
    def __dlpack_device__(self):
        return (kDLROCM, self.device)

    def __dlpack__(self, stream=None):
        """Exports a view. Work queued by this library on its own stream is
        complete first (stream-ordered consumers then see the data)."""
        _lib.check(_lib.load().npph_stream_sync(None))
        return to_dlpack(self)

    @classmethod
    def from_dlpack(cls, obj):
        """A zero-copy view of a C-contiguous ROCm DLPack tensor (an object with
        __dlpack__, e.g. a torch tensor on a ROCm device, or a capsule)."""
        cap = obj.__dlpack__() if hasattr(obj, "__dlpack__") else obj
        p = _capi.PyCapsule_GetPointer(cap, b"dltensor")
        mt = ctypes.cast(p, ctypes.POINTER(DLManagedTensor))
        t = mt.contents.dl_tensor
        if t.device.device_type != kDLROCM:
            raise TypeError(f"DLPack device type {t.device.device_type} is not ROCm ({kDLROCM})")
        shape = tuple(t.shape[i] for i in range(t.ndim))
        if t.strides:
            expect, acc = [], 1
            for d in reversed(shape):
                expect.insert(0, acc)
                acc *= d
            if any(shape[i] > 1 and t.strides[i] != expect[i] for i in range(t.ndim)):
                raise ValueError("DLPack tensor is not C-contiguous")
        _capi.PyCapsule_SetName(cap, b"used_dltensor")
        owner = _Imported(mt)
        arr = cls(shape, _np_dtype(t.dtype), _ptr=(t.data or 0) + t.byte_offset, _owner=owner)
        arr.device = t.device.device_id
        return arr

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
        if self._owner is not None:
            self._owner.release()
            self._owner = None
            self.ptr = None
        elif self.ptr:
            _lib.load().npph_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _dl_dtype(dt: np.dtype) -> DLDataType:
    return DLDataType(_DL_CODES[dt.kind], dt.itemsize * 8, 1)


def _np_dtype(d: DLDataType) -> np.dtype:
    kinds = {v: k for k, v in _DL_CODES.items()}
    if d.lanes != 1 or d.code not in kinds:
        raise TypeError(f"DLPack dtype code {d.code} / {d.bits} bits / {d.lanes} lanes is not supported")
    return np.dtype(f"{kinds[d.code]}{d.bits // 8}") if d.code != 6 else np.dtype(bool)


def to_dlpack(arr: "DeviceArray"):
    """A "dltensor" PyCapsule viewing arr's device memory (kDLROCM, C order).
    arr stays alive until the consumer calls the tensor's deleter."""
    mt = DLManagedTensor()
    shape = (ctypes.c_int64 * max(1, len(arr.shape)))(*arr.shape)
    mt.dl_tensor.data = arr.ptr.value if arr.ptr else None
    mt.dl_tensor.device = DLDevice(kDLROCM, arr.device)
    mt.dl_tensor.ndim = len(arr.shape)
    mt.dl_tensor.dtype = _dl_dtype(arr.dtype)
    mt.dl_tensor.shape = ctypes.cast(shape, ctypes.POINTER(ctypes.c_int64))
    mt.dl_tensor.strides = None
    mt.dl_tensor.byte_offset = 0
    mt.manager_ctx = None
    mt.deleter = _dl_deleter
    addr = ctypes.addressof(mt)
    _exported[addr] = (mt, shape, arr)
    return _capi.PyCapsule_New(addr, b"dltensor", ctypes.cast(_capsule_dtor, ctypes.c_void_p))


class _Imported:
    """Keeps a consumed DLManagedTensor until the view is released."""

    def __init__(self, mt_ptr):
        self.mt = mt_ptr

    def release(self):
        if self.mt:
            m = self.mt.contents
            if m.deleter:
                m.deleter(self.mt)
            self.mt = None


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
