"""Minimal HIP runtime access through ctypes: device buffers, copies, streams
and events.  It binds the same ``libamdhip64`` that ``libfftoc.so`` links, so
buffers and streams created here are valid arguments to the library.
"""
from __future__ import annotations

import ctypes

import numpy as np

_hip = None

hipMemcpyHostToDevice = 1
hipMemcpyDeviceToHost = 2
hipMemcpyDeviceToDevice = 3


class HipError(RuntimeError):
    pass


def lib():
    global _hip
    if _hip is None:
        from . import _native  # loads libfftoc.so first, which pulls in libamdhip64
        _native.load()
        h = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        vp = ctypes.c_void_p
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        h.hipMemcpyAsync.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, vp]
        h.hipMemsetAsync.argtypes = [vp, ctypes.c_int, ctypes.c_size_t, vp]
        h.hipDeviceSynchronize.argtypes = []
        h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipStreamDestroy.argtypes = [vp]
        h.hipStreamSynchronize.argtypes = [vp]
        h.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipEventDestroy.argtypes = [vp]
        h.hipEventRecord.argtypes = [vp, vp]
        h.hipEventSynchronize.argtypes = [vp]
        h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        h.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        h.hipGetDevice.argtypes = [ctypes.POINTER(ctypes.c_int)]
        h.hipDeviceGetAttribute.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int]
        h.hipGetErrorString.argtypes = [ctypes.c_int]
        h.hipGetErrorString.restype = ctypes.c_char_p
        h.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        _hip = h
    return _hip


def check(rc: int, what: str = "HIP call") -> None:
    if rc != 0:
        msg = lib().hipGetErrorString(rc).decode()
        raise HipError(f"{what} failed: {msg} ({rc})")


def device_count() -> int:
    n = ctypes.c_int(0)
    rc = lib().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


_ATTR_MULTIPROCESSOR_COUNT = 63  # hipDeviceAttributeMultiprocessorCount (hip_runtime_api.h)


def cu_count() -> int:
    """Compute units of the current device."""
    dev, v = ctypes.c_int(0), ctypes.c_int(0)
    check(lib().hipGetDevice(ctypes.byref(dev)), "hipGetDevice")
    check(lib().hipDeviceGetAttribute(ctypes.byref(v), _ATTR_MULTIPROCESSOR_COUNT, dev.value),
          "hipDeviceGetAttribute")
    return v.value


def synchronize() -> None:
    """Waits for the legacy default stream (and the blocking streams it
    orders with).  Unlike hipDeviceSynchronize it does not wait for work on
    non-blocking streams, such as a running server loop."""
    check(lib().hipStreamSynchronize(None), "hipStreamSynchronize")


def device_synchronize() -> None:
    check(lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    check(lib().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value, total.value


def download_ptr(ptr: int, shape, dtype=np.complex128) -> np.ndarray:
    """Copies device memory at a raw address into a new host array."""
    out = np.empty(shape, dtype=dtype)
    check(lib().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), out.nbytes,
                          hipMemcpyDeviceToHost), "hipMemcpy D2H")
    return out


class DeviceBuffer:
    """Owning device allocation; ``ptr`` is the raw device address."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), f"hipMalloc({self.nbytes})")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        check(lib().hipMemcpy(ctypes.c_void_p(self.ptr + offset), a.ctypes.data_as(ctypes.c_void_p),
                              a.nbytes, hipMemcpyHostToDevice), "hipMemcpy H2D")

    def download(self, shape, dtype=np.complex128, offset: int = 0) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        check(lib().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr + offset),
                              out.nbytes, hipMemcpyDeviceToHost), "hipMemcpy D2H")
        return out

    def copy_from(self, other: "DeviceBuffer", nbytes: int | None = None) -> None:
        n = self.nbytes if nbytes is None else nbytes
        check(lib().hipMemcpy(ctypes.c_void_p(self.ptr), ctypes.c_void_p(other.ptr), n,
                              hipMemcpyDeviceToDevice), "hipMemcpy D2D")

    def free(self) -> None:
        if self.ptr:
            lib().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        check(lib().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self) -> None:
        check(lib().hipStreamSynchronize(ctypes.c_void_p(self.handle)), "hipStreamSynchronize")

    def __del__(self):
        try:
            if self.handle:
                lib().hipStreamDestroy(ctypes.c_void_p(self.handle))
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(lib().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream: "Stream | None" = None) -> None:
        s = ctypes.c_void_p(stream.handle if stream else None)
        check(lib().hipEventRecord(ctypes.c_void_p(self.handle), s), "hipEventRecord")

    def synchronize(self) -> None:
        check(lib().hipEventSynchronize(ctypes.c_void_p(self.handle)), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0)
        check(lib().hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(self.handle),
                                        ctypes.c_void_p(end.handle)), "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                lib().hipEventDestroy(ctypes.c_void_p(self.handle))
        except Exception:
            pass
