"""Minimal ctypes binding to the HIP runtime that libfdmgpu.so itself uses.

Device buffers, streams and events for tests, bench.py and callers that do
not bring their own device memory.  Loading libfdmgpu.so first pins the
process to /opt/rocm's libamdhip64.so.7, and this module binds to that same
already-loaded library, so all device pointers belong to one runtime.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as _fdm_lib

_hip = None
H2D, D2H, D2D = 1, 2, 3


def rt():
    global _hip
    if _hip is None:
        _fdm_lib()   # loads libamdhip64.so.7 as a dependency of libfdmgpu.so
        H = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        vp, sz, i = C.c_void_p, C.c_size_t, C.c_int
        H.hipMalloc.argtypes = [C.POINTER(vp), sz]
        H.hipFree.argtypes = [vp]
        H.hipMemcpy.argtypes = [vp, vp, sz, i]
        H.hipMemcpyAsync.argtypes = [vp, vp, sz, i, vp]
        H.hipMemsetAsync.argtypes = [vp, i, sz, vp]
        H.hipStreamCreate.argtypes = [C.POINTER(vp)]
        H.hipStreamDestroy.argtypes = [vp]
        H.hipStreamSynchronize.argtypes = [vp]
        H.hipDeviceSynchronize.argtypes = []
        H.hipEventCreate.argtypes = [C.POINTER(vp)]
        H.hipEventDestroy.argtypes = [vp]
        H.hipEventRecord.argtypes = [vp, vp]
        H.hipEventSynchronize.argtypes = [vp]
        H.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), vp, vp]
        H.hipGetErrorString.argtypes = [i]
        H.hipGetErrorString.restype = C.c_char_p
        H.hipGetDeviceCount.argtypes = [C.POINTER(i)]
        H.hipMemGetInfo.argtypes = [C.POINTER(sz), C.POINTER(sz)]
        _hip = H
    return _hip


def check(rc: int, what: str = "hip"):
    if rc != 0:
        raise RuntimeError(f"{what}: {rt().hipGetErrorString(rc).decode()} ({rc})")


def device_count() -> int:
    n = C.c_int(0)
    rc = rt().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def mem_info():
    free, total = C.c_size_t(), C.c_size_t()
    check(rt().hipMemGetInfo(C.byref(free), C.byref(total)), "hipMemGetInfo")
    return free.value, total.value


def synchronize():
    check(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    def __init__(self):
        self.h = C.c_void_p()
        check(rt().hipStreamCreate(C.byref(self.h)), "hipStreamCreate")

    def synchronize(self):
        check(rt().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _hip is not None:
            _hip.hipStreamDestroy(self.h)


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        check(rt().hipEventCreate(C.byref(self.h)), "hipEventCreate")

    def record(self, stream: Stream | None = None):
        check(rt().hipEventRecord(self.h, stream.h if stream else None), "hipEventRecord")

    def synchronize(self):
        check(rt().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(rt().hipEventElapsedTime(C.byref(ms), self.h, end.h), "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _hip is not None:
            _hip.hipEventDestroy(self.h)


class DeviceArray:
    """A device allocation with numpy-like shape/dtype and host transfers."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        check(rt().hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        check(rt().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes, H2D), "hipMemcpy H2D")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(rt().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def offset(self, elems: int) -> C.c_void_p:
        return C.c_void_p(self.ptr.value + elems * self.dtype.itemsize)

    def free(self):
        if self.ptr.value and _hip is not None:
            _hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        self.free()
