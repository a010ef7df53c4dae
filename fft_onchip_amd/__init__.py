"""fft_onchip_amd -- batched small complex FFTs (N = 8..4096) for MI355X.

The product is the C-ABI library ``libfftoc.so`` (include/fftoc.h) with
hand-written gfx950 HIP kernels.  This module is its host-side mirror: the
same entry points as the reference's size-templated functors
(``tensor_fft_N``, ``legacy_fft`` = radix 8, ``legacy16_fft`` = radix 16;
includes/tensor_fft_*.cuh, includes/legacy8_fft.cuh, includes/legacy16_fft.cuh)
taking device buffers, plus a host-array convenience wrapper.

Data layout is the reference's: interleaved complex<double>, each signal
contiguous, signals back to back (includes/config.hpp:6).  Transforms run in
place when ``out`` is omitted, like the reference's in-shared-memory functors.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .hip import DeviceBuffer, Event, Stream, device_count, synchronize  # noqa: F401

FORWARD = 0x0
INVERSE = 0x1
NORMALIZE = 0x2
PREC_FP64 = 0x00
PREC_FP32 = 0x10
PREC_BF16X3 = 0x20
DATA_C64 = 0x100
_PREC = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3}

STATUS = {
    0: "SUCCESS", 1: "INVALID_SIZE", 2: "INVALID_BATCH", 3: "NULL_POINTER", 4: "INVALID_FLAGS",
    5: "OVERLAP", 6: "UNSUPPORTED", 7: "DEVICE", 8: "LAUNCH", 9: "ALLOC", 10: "MISALIGNED",
    11: "BUDGET", 12: "QUEUE_FULL", 13: "INVALID_HANDLE",
}
TENSOR_SIZES = _native.TENSOR_SIZES
LEGACY8_SIZES = (8, 64, 512, 4096)
LEGACY16_SIZES = (16, 256, 4096)
ALL_SIZES = tuple(1 << k for k in range(3, 13))


class FftocError(RuntimeError):
    def __init__(self, status: int, where: str = ""):
        self.status = status
        self.name = STATUS.get(status, str(status))
        msg = _native.load().fftoc_strerror(status).decode()
        super().__init__(f"{where}: FFTOC_ERROR_{self.name} ({msg})" if where else msg)


def lib():
    return _native.load()


def _ptr(x) -> int | None:
    if x is None:
        return None
    if isinstance(x, DeviceBuffer):
        return x.ptr
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return int(x.data_ptr())
    raise TypeError(f"cannot take a device pointer from {type(x)!r}")


def _stream(s) -> int | None:
    if s is None:
        return None
    if isinstance(s, Stream):
        return s.handle
    if isinstance(s, int):
        return s
    if hasattr(s, "cuda_stream"):
        return int(s.cuda_stream)
    raise TypeError(f"cannot take a stream handle from {type(s)!r}")


def make_flags(inverse: bool = False, normalize: bool = False, precision: str = "fp64",
               tile: int = 0, per_cu: int = 0) -> int:
    """Flag bits: direction, 1/N scaling, stage precision, FFTOC_TILE(tile)
    and FFTOC_WGS_PER_CU(per_cu) (0 = the tuned choice)."""
    f = (INVERSE if inverse else 0) | (NORMALIZE if normalize else 0) | _PREC[precision]
    return f | ((int(tile) & 0xFF) << 16) | ((int(per_cu) & 0x3F) << 24)


def _call(fn_name: str, x, out, batch: int, n: int | None, flags: int, stream) -> None:
    f = getattr(lib(), fn_name)
    ip = _ptr(x)
    op = _ptr(out) if out is not None else ip
    if n is None:
        rc = f(ip, op, int(batch), int(flags), _stream(stream))
    else:
        rc = f(ip, op, int(batch), int(n), int(flags), _stream(stream))
    if rc != 0:
        raise FftocError(rc, fn_name)


def tensor_fft(x, batch: int, n: int, out=None, flags: int = 0, stream=None) -> None:
    """FP64-MFMA Karatsuba path (the reference's tensor_fft_N functors)."""
    _call("fftoc_tensor_fft", x, out, batch, n, flags, stream)


def _make_sized(n):
    def f(x, batch: int, out=None, flags: int = 0, stream=None) -> None:
        _call(f"fftoc_tensor_fft_{n}", x, out, batch, None, flags, stream)
    f.__name__ = f"tensor_fft_{n}"
    f.__doc__ = f"{n}-point FP64 MFMA FFT (includes/tensor_fft_{n}.cuh)."
    return f


for _n in TENSOR_SIZES:
    globals()[f"tensor_fft_{_n}"] = _make_sized(_n)


def legacy_fft(x, batch: int, n: int, out=None, flags: int = 0, stream=None) -> None:
    """Radix-8 register-butterfly path (includes/legacy8_fft.cuh), n = 8^k."""
    _call("fftoc_legacy8_fft", x, out, batch, n, flags, stream)


def legacy16_fft(x, batch: int, n: int, out=None, flags: int = 0, stream=None) -> None:
    """Radix-16 register-butterfly path (includes/legacy16_fft.cuh), n = 16^k."""
    _call("fftoc_legacy16_fft", x, out, batch, n, flags, stream)


def fft(x, batch: int, n: int, out=None, flags: int = 0, stream=None) -> None:
    """Fastest FP64 kernel for n (any power of two 8..4096)."""
    _call("fftoc_fft", x, out, batch, n, flags, stream)


def variants(n: int, c64: bool = False) -> list:
    """Names of the kernel variants of length n for complex128 (or, with
    c64, complex64) data, in index order."""
    dt = DATA_C64 if c64 else 0
    return [lib().fftoc_variant_name(int(n), dt, v).decode()
            for v in range(lib().fftoc_variant_count(int(n), dt))]


def variant_index(n: int, name: str, c64: bool = False) -> int:
    """Index of the kernel variant called `name` for length n and data type."""
    try:
        return variants(n, c64).index(name)
    except ValueError:
        raise KeyError(f"no variant {name!r} for n={n}, c64={c64}") from None


def fft_variant(x, batch: int, n: int, variant: int, out=None, flags: int = 0,
                stream=None) -> None:
    """One named kernel variant (fftoc_fft_variant)."""
    ip = _ptr(x)
    op = _ptr(out) if out is not None else ip
    rc = lib().fftoc_fft_variant(ip, op, int(batch), int(n), int(flags), int(variant),
                                 _stream(stream))
    if rc != 0:
        raise FftocError(rc, "fftoc_fft_variant")


def variant_occupancy(n: int, variant: int, c64: bool = False) -> int:
    """Resident workgroups per CU of a variant's kernel on the current device
    (the upper end of its workgroups-per-CU range)."""
    return int(lib().fftoc_variant_occupancy(int(n), DATA_C64 if c64 else 0, int(variant)))


def workgroups_per_cu(n: int, batch: int, c64: bool = False) -> int:
    """Workgroups-per-CU cap fftoc_fft uses at (n, batch bucket); 0 = the
    kernel's occupancy."""
    return int(lib().fftoc_get_workgroups_per_cu(int(n), DATA_C64 if c64 else 0, int(batch)))


def set_workgroups_per_cu(n: int, batch: int, per_cu: int, c64: bool = False) -> None:
    """Overrides that cap (batch < 0: every batch bucket)."""
    rc = lib().fftoc_set_workgroups_per_cu(int(n), DATA_C64 if c64 else 0, int(batch), int(per_cu))
    if rc != 0:
        raise FftocError(rc, "fftoc_set_workgroups_per_cu")


_PATHS = {"tensor": tensor_fft, "legacy8": legacy_fft, "legacy16": legacy16_fft, "best": fft}


def fft_host(x: np.ndarray, path: str = "best", inverse: bool = False, normalize: bool = False,
             precision: str = "fp64", in_place: bool = False) -> np.ndarray:
    """Uploads signals (batch, n) -- complex128, or complex64 for FP32 data --
    transforms them on the GPU and returns the result as a new host array."""
    x = np.asarray(x)
    if x.dtype != np.complex64:
        x = x.astype(np.complex128)
    x = np.ascontiguousarray(x)
    if x.ndim == 1:
        x = x.reshape(1, -1)
    batch, n = x.shape
    flags = make_flags(inverse, normalize, precision) | (DATA_C64 if x.dtype == np.complex64 else 0)
    src = DeviceBuffer.from_array(x)
    if in_place:
        _PATHS[path](src, batch, n, None, flags)
        res = src
    else:
        res = DeviceBuffer(x.nbytes)
        _PATHS[path](src, batch, n, res, flags)
    synchronize()
    return res.download(x.shape, dtype=x.dtype)


class Plan:
    """Mixed-size batch plan (fftoc_plan_*): one persistent kernel launch over
    many jobs.  jobs: iterable of (in_buf, out_buf, batch, n, flags)."""

    def __init__(self, jobs):
        self._jobs = list(jobs)
        arr = (_native.Job * max(1, len(self._jobs)))()
        for i, (src, dst, batch, n, flags) in enumerate(self._jobs):
            arr[i] = _native.Job(_ptr(src), _ptr(dst) if dst is not None else _ptr(src),
                                 int(batch), int(n), int(flags))
        handle = ctypes.c_void_p()
        rc = lib().fftoc_plan_create(arr, len(self._jobs), ctypes.byref(handle))
        if rc != 0:
            raise FftocError(rc, "fftoc_plan_create")
        self.handle = handle.value

    def execute(self, stream=None, graph: bool = False) -> None:
        fn = lib().fftoc_plan_execute_graph if graph else lib().fftoc_plan_execute
        rc = fn(self.handle, _stream(stream))
        if rc != 0:
            raise FftocError(rc, "fftoc_plan_execute")

    def autotune(self, reps: int = 3, stream=None) -> None:
        rc = lib().fftoc_plan_autotune(self.handle, int(reps), _stream(stream))
        if rc != 0:
            raise FftocError(rc, "fftoc_plan_autotune")

    def routine(self, n: int, c64: bool = False) -> str:
        name = ctypes.c_char_p()
        rc = lib().fftoc_plan_routine(self.handle, int(n), DATA_C64 if c64 else 0, ctypes.byref(name))
        if rc != 0:
            raise FftocError(rc, "fftoc_plan_routine")
        return name.value.decode()

    def set_routine(self, n: int, name: str, c64: bool = False) -> None:
        """Pins the tile routine of length n (one of batch_routines(n))."""
        rc = lib().fftoc_plan_set_routine(self.handle, int(n), DATA_C64 if c64 else 0,
                                          name.encode())
        if rc != 0:
            raise FftocError(rc, "fftoc_plan_set_routine")

    @property
    def workgroups_per_cu(self) -> int:
        """Workgroups per CU of the plan's launch (0 = the default)."""
        return int(lib().fftoc_plan_workgroups_per_cu(self.handle))

    @workgroups_per_cu.setter
    def workgroups_per_cu(self, k: int) -> None:
        rc = lib().fftoc_plan_set_workgroups_per_cu(self.handle, int(k))
        if rc != 0:
            raise FftocError(rc, "fftoc_plan_set_workgroups_per_cu")

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().fftoc_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


NOALLOC_MAX_JOBS = 32


def batch_fft_noalloc(jobs, stream=None) -> None:
    """One launch over up to NOALLOC_MAX_JOBS jobs (in_buf, out_buf, batch, n,
    flags) that allocates nothing (fftoc_batch_fft_noalloc): same results as
    a Plan of the same jobs.  Calls overlapping on one device race."""
    jobs = list(jobs)
    arr = (_native.Job * max(1, len(jobs)))()
    for i, (src, dst, batch, n, flags) in enumerate(jobs):
        arr[i] = _native.Job(_ptr(src), _ptr(dst) if dst is not None else _ptr(src),
                             int(batch), int(n), int(flags))
    rc = lib().fftoc_batch_fft_noalloc(arr, len(jobs), _stream(stream))
    if rc != 0:
        raise FftocError(rc, "fftoc_batch_fft_noalloc")


def batch_routines(n: int) -> list:
    """Names of the tile routines plans and servers can use for length n."""
    names, i = [], 0
    while True:
        name = ctypes.c_char_p()
        rc = lib().fftoc_batch_routine_name(int(n), i, ctypes.byref(name))
        if rc == 13:  # FFTOC_ERROR_INVALID_HANDLE: past the last routine
            return names
        if rc != 0:
            raise FftocError(rc, "fftoc_batch_routine_name")
        names.append(name.value.decode())
        i += 1


class Server:
    """Device-resident FFT server (fftoc_server_*)."""

    def __init__(self, budget_bytes: int, max_inflight: int = 256):
        h = ctypes.c_void_p()
        rc = lib().fftoc_server_create(int(budget_bytes), int(max_inflight), ctypes.byref(h))
        if rc != 0:
            raise FftocError(rc, "fftoc_server_create")
        self.handle = h.value

    def submit(self, x, batch: int, n: int, flags: int = 0):
        """Returns the request id, or None if no slot / budget is free now."""
        rid = ctypes.c_uint64()
        rc = lib().fftoc_server_submit(self.handle, _ptr(x), int(batch), int(n), int(flags),
                                       ctypes.byref(rid))
        if rc == 12:  # FFTOC_ERROR_QUEUE_FULL
            return None
        if rc != 0:
            raise FftocError(rc, "fftoc_server_submit")
        return rid.value

    def poll(self, rid: int):
        """Device pointer of the result when done, else None."""
        done, out = ctypes.c_int(0), ctypes.c_void_p()
        rc = lib().fftoc_server_poll(self.handle, ctypes.c_uint64(rid), ctypes.byref(done), ctypes.byref(out))
        if rc != 0:
            raise FftocError(rc, "fftoc_server_poll")
        return out.value if done.value else None

    def wait(self, rid: int, timeout_s: float = 20.0):
        import time
        t0 = time.perf_counter()
        while True:
            p = self.poll(rid)
            if p is not None:
                return p
            if time.perf_counter() - t0 > timeout_s:
                raise TimeoutError(f"request {rid} not complete after {timeout_s} s")

    def release(self, rid: int) -> None:
        rc = lib().fftoc_server_release(self.handle, ctypes.c_uint64(rid))
        if rc != 0:
            raise FftocError(rc, "fftoc_server_release")

    def set_idle_timeout(self, seconds: float) -> None:
        rc = lib().fftoc_server_set_idle_timeout(self.handle, float(seconds))
        if rc != 0:
            raise FftocError(rc, "fftoc_server_set_idle_timeout")

    def stats(self):
        used, inflight = ctypes.c_size_t(0), ctypes.c_int(0)
        lib().fftoc_server_stats(self.handle, ctypes.byref(used), ctypes.byref(inflight))
        return used.value, inflight.value

    def replay(self, reqs, arrivals):
        """reqs: list of (in_buf, batch, n, flags); arrivals: seconds.
        Returns (latencies_s, span_s)."""
        count = len(reqs)
        arr = (_native.Job * max(1, count))()
        for i, (src, batch, n, flags) in enumerate(reqs):
            arr[i] = _native.Job(_ptr(src), None, int(batch), int(n), int(flags))
        at = (ctypes.c_double * max(1, count))(*arrivals)
        lat = (ctypes.c_double * max(1, count))()
        span = ctypes.c_double(0)
        rc = lib().fftoc_server_replay(self.handle, arr, count, at, lat, ctypes.byref(span))
        if rc != 0:
            raise FftocError(rc, "fftoc_server_replay")
        return np.array(lat[:count]), span.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            rc = lib().fftoc_server_destroy(self.handle)
            self.handle = None
            if rc != 0:
                raise FftocError(rc, "fftoc_server_destroy")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def init() -> None:
    rc = lib().fftoc_init()
    if rc != 0:
        raise FftocError(rc, "fftoc_init")


def version() -> int:
    return int(lib().fftoc_version())
