"""HBM buffers, streams and events for Python harnesses, on top of the C ABI.

No torch: device memory comes from the library's own sx_malloc so that a
single HIP runtime (the one the library links) owns every pointer.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class DeviceArray:
    """A typed device allocation.  Freed when garbage collected."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        # at least 16 bytes so that empty arrays still have a valid address
        _lib.call("sx_malloc", C.byref(p), max(self.nbytes, 16))
        self.ptr = p.value
        self._lib = _lib.load()

    @classmethod
    def from_numpy(cls, arr: np.ndarray, stream=None) -> "DeviceArray":
        arr = np.ascontiguousarray(arr)
        d = cls(arr.shape, arr.dtype)
        if d.nbytes:
            _lib.call("sx_memcpy_htod", d.ptr, arr.ctypes.data, d.nbytes, stream)
            _lib.call("sx_stream_sync", stream)
        return d

    @classmethod
    def zeros(cls, shape, dtype, stream=None) -> "DeviceArray":
        d = cls(shape, dtype)
        _lib.call("sx_fill_u8", d.ptr, 0, d.nbytes, stream)
        _lib.call("sx_stream_sync", stream)
        return d

    def numpy(self, stream=None) -> np.ndarray:
        """Host copy.  With a stream, the copy follows that stream's work.
        Without one, every stream's work is waited for first: the library's
        streams are non-blocking, so the null stream alone would not order
        the copy after a kernel launched on one of them."""
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            if stream is None:
                _lib.call("sx_device_sync")
            _lib.call("sx_memcpy_dtoh", out.ctypes.data, self.ptr, self.nbytes, stream)
        return out

    def fill_(self, byte: int, stream=None) -> "DeviceArray":
        _lib.call("sx_fill_u8", self.ptr, byte, self.nbytes, stream)
        return self

    def offset(self, elements: int) -> int:
        """Device address of element `elements` (flat)."""
        return self.ptr + int(elements) * self.dtype.itemsize

    def free(self) -> None:
        if getattr(self, "ptr", None):
            self._lib.sx_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __len__(self):
        return self.shape[0]

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype}, ptr={self.ptr:#x})"


def fetch(ptr: int, count: int, dtype, stream=None) -> np.ndarray:
    """Copy `count` elements of `dtype` from a raw device pointer (without a
    stream, after every stream's work, as DeviceArray.numpy)."""
    dt = np.dtype(dtype)
    out = np.empty(int(count), dt)
    if count:
        if stream is None:
            _lib.call("sx_device_sync")
        _lib.call("sx_memcpy_dtoh", out.ctypes.data, ptr, out.nbytes, stream)
    return out


class Stream:
    def __init__(self):
        p = C.c_void_p()
        _lib.call("sx_stream_create", C.byref(p))
        self.handle = p.value
        self._lib = _lib.load()

    def sync(self) -> None:
        _lib.call("sx_stream_sync", self.handle)

    def __del__(self):
        try:
            if self.handle:
                self._lib.sx_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        p = C.c_void_p()
        _lib.call("sx_event_create", C.byref(p))
        self.handle = p.value
        self._lib = _lib.load()

    def record(self, stream=None) -> "Event":
        _lib.call("sx_event_record", self.handle, stream)
        return self

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        _lib.call("sx_event_elapsed_ms", self.handle, end.handle, C.byref(ms))
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                self._lib.sx_event_destroy(self.handle)
        except Exception:
            pass


def device_count() -> int:
    n = C.c_int(0)
    rc = _lib.load().sx_device_count(C.byref(n))
    return n.value if rc == _lib.SX_OK else 0


def device_name(dev: int = 0) -> str:
    buf = C.create_string_buffer(256)
    _lib.call("sx_device_name", dev, buf, 256)
    return buf.value.decode()


def synchronize() -> None:
    _lib.call("sx_device_sync")
