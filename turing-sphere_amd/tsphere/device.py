"""Device memory, streams and events through the C ABI.

The reference allocates WebGPU buffers (src/mgpu/plan.ts:33-38) and orders
work on one queue; here a :class:`DeviceBuffer` is a hipMalloc'd float32
array and a :class:`Stream` a HIP stream at default priority.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def device_count() -> int:
    n = C.c_int(0)
    _lib.call("tsphere_device_count", C.byref(n))
    return n.value


class Stream:
    """A HIP stream (default priority).  ``Stream.null()`` is stream 0."""

    def __init__(self, handle=None, owned: bool = True):
        if handle is None and owned:
            h = C.c_void_p()
            _lib.call("tsphere_stream_create", C.byref(h))
            handle = h.value
        self.handle = handle
        self._owned = owned and handle is not None

    @staticmethod
    def null() -> "Stream":
        return Stream(None, owned=False)

    def sync(self) -> None:
        _lib.call("tsphere_stream_sync", self.handle)

    def destroy(self) -> None:
        if self._owned and self.handle:
            _lib.call("tsphere_stream_destroy", self.handle)
        self.handle = None
        self._owned = False

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class DeviceBuffer:
    """``count`` float32 values in device memory (caller-owned)."""

    def __init__(self, count: int, label: str = ""):
        self.count = int(count)
        self.label = label
        self.nbytes = 4 * self.count
        p = C.c_void_p()
        _lib.call("tsphere_malloc", C.c_size_t(max(self.nbytes, 4)), C.byref(p))
        self.ptr = p.value
        self._view_of = None

    @classmethod
    def view(cls, parent: "DeviceBuffer", offset: int, count: int, label: str = "") -> "DeviceBuffer":
        """A sub-range of ``parent`` (offset/count in floats), not owning memory."""
        if offset < 0 or offset + count > parent.count:
            raise ValueError("view out of range")
        b = cls.__new__(cls)
        b.count = int(count)
        b.label = label
        b.nbytes = 4 * b.count
        b.ptr = parent.ptr + 4 * offset
        b._view_of = parent
        return b

    def upload(self, data, stream: Stream | None = None) -> None:
        a = np.ascontiguousarray(data, np.float32).reshape(-1)
        if a.size != self.count:
            raise ValueError(f"upload '{self.label}': expected {self.count} elements, got {a.size}")
        s = stream.handle if stream else None
        _lib.call("tsphere_memcpy_htod", self.ptr, a.ctypes.data, C.c_size_t(self.nbytes), s)
        # keep the host array alive until the copy has been consumed
        _lib.call("tsphere_stream_sync", s)

    def download(self, stream: Stream | None = None) -> np.ndarray:
        out = np.empty(self.count, np.float32)
        s = stream.handle if stream else None
        _lib.call("tsphere_memcpy_dtoh", out.ctypes.data, self.ptr, C.c_size_t(self.nbytes), s)
        return out

    def copy_from(self, other: "DeviceBuffer", stream: Stream | None = None) -> None:
        if other.count != self.count:
            raise ValueError("copy_from: size mismatch")
        s = stream.handle if stream else None
        _lib.call("tsphere_memcpy_dtod", self.ptr, other.ptr, C.c_size_t(self.nbytes), s)

    def fill(self, value: float, stream: Stream | None = None) -> None:
        s = stream.handle if stream else None
        _lib.call("tsphere_fill_f32", self.ptr, C.c_float(value), C.c_longlong(self.count), s)

    def free(self) -> None:
        if self.ptr and self._view_of is None:
            _lib.call("tsphere_free", self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        h = C.c_void_p()
        _lib.call("tsphere_event_create", C.byref(h))
        self.handle = h.value

    def record(self, stream: Stream | None = None) -> None:
        _lib.call("tsphere_event_record", self.handle, stream.handle if stream else None)

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        _lib.call("tsphere_event_elapsed", self.handle, end.handle, C.byref(ms))
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                _lib.call("tsphere_event_destroy", self.handle)
        except Exception:
            pass
