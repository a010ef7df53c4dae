"""CPU oracle for the batched small complex FFT -- TEST INFRASTRUCTURE ONLY.

Only the test suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or, for the
baseline, as the timed CPU transform).  The product library never calls it.

The C restatement lives in ``fftoc_oracle.c`` (built to ``liboracle.so`` by
``make -C oracle``); this module is a thin ctypes wrapper plus the error
metrics the tests use.  Signals are complex128 arrays of shape (batch, n),
contiguous interleaved (re, im) like ``thrust::complex<double>`` in the
reference (includes/config.hpp:6).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i64 = ctypes.c_int64
_dp = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> None:
    """Compile the C oracle with gcc (``make -C oracle``)."""
    cmd = ["make", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        for name in ("oracle_dft", "oracle_fft"):
            f = getattr(lib, name)
            f.argtypes = [_i64, _i64, _dp, _dp, _dp, ctypes.c_int]
            f.restype = ctypes.c_int
        lib.oracle_fft_f64.argtypes = [_i64, _i64, _dp, _dp, ctypes.c_int, ctypes.c_int]
        lib.oracle_fft_f64.restype = ctypes.c_int
        for name in ("oracle_legacy8", "oracle_legacy16"):
            f = getattr(lib, name)
            f.argtypes = [_i64, _i64, _dp, _dp, _dp]
            f.restype = ctypes.c_int
        lib.oracle_karatsuba.argtypes = [_i64, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                         _i64, _dp, _dp, _dp]
        lib.oracle_karatsuba.restype = ctypes.c_int
        lib.oracle_twiddles.argtypes = [_i64, ctypes.c_int, _dp, _dp]
        lib.oracle_twiddles.restype = ctypes.c_int
        _lib = lib
    return _lib


def _as_signals(x, n=None):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    if x.ndim == 1:
        x = x.reshape(1, -1)
    if n is not None and x.shape[1] != n:
        raise ValueError(f"expected signals of length {n}, got {x.shape}")
    return x


def _dptr(a):
    return a.ctypes.data_as(_dp)


class DD:
    """Double-double result: ``hi + lo`` carries the oracle's long-double value."""

    def __init__(self, hi: np.ndarray, lo: np.ndarray):
        self.hi = hi
        self.lo = lo

    def as_longdouble(self) -> np.ndarray:
        return self.hi.astype(np.clongdouble) + self.lo.astype(np.clongdouble)


def _run_dd(fn_name, x, *extra):
    lib = _load()
    x = _as_signals(x)
    b, n = x.shape
    hi = np.empty_like(x)
    lo = np.empty_like(x)
    rc = getattr(lib, fn_name)(n, b, _dptr(x.view(np.float64)), _dptr(hi.view(np.float64)),
                               _dptr(lo.view(np.float64)), *extra)
    if rc != 0:
        raise ValueError(f"{fn_name}(n={n}) failed with code {rc}")
    return DD(hi, lo)


def dft(x, inverse: bool = False) -> DD:
    """Brute-force O(N^2) long-double DFT (unnormalised in both directions)."""
    return _run_dd("oracle_dft", x, int(inverse))


def fft(x, inverse: bool = False) -> DD:
    """Radix-2 long-double FFT; n must be a power of two."""
    return _run_dd("oracle_fft", x, int(inverse))


def legacy8(x) -> DD:
    """Radix-8 Stockham restatement of includes/legacy8_fft.cuh:57-118."""
    return _run_dd("oracle_legacy8", x)


def legacy16(x) -> DD:
    """Radix-16 Stockham restatement of includes/legacy16_fft.cuh:99-172."""
    return _run_dd("oracle_legacy16", x)


def karatsuba(x, radices) -> DD:
    """DFT-as-matrix stages with 3-real-product complex GEMMs
    (includes/c64_fft64.cuh:8-27, includes/tensor_utils.cuh:25-36)."""
    lib = _load()
    x = _as_signals(x)
    b, n = x.shape
    r = (ctypes.c_int * len(radices))(*radices)
    hi = np.empty_like(x)
    lo = np.empty_like(x)
    rc = lib.oracle_karatsuba(n, r, len(radices), b, _dptr(x.view(np.float64)),
                              _dptr(hi.view(np.float64)), _dptr(lo.view(np.float64)))
    if rc != 0:
        raise ValueError(f"oracle_karatsuba(n={n}, radices={radices}) failed with code {rc}")
    return DD(hi, lo)


def fft_f64(x, inverse: bool = False, threads: int = 1) -> np.ndarray:
    """Float64 Stockham FFT on the host (the timed CPU baseline)."""
    lib = _load()
    x = _as_signals(x)
    b, n = x.shape
    out = np.empty_like(x)
    rc = lib.oracle_fft_f64(n, b, _dptr(x.view(np.float64)), _dptr(out.view(np.float64)),
                            int(inverse), int(threads))
    if rc != 0:
        raise ValueError(f"oracle_fft_f64(n={n}) failed with code {rc}")
    return out


def twiddles(n: int, inverse: bool = False) -> DD:
    """exp(-+2 pi i j / n), j = 0..n-1, from the oracle's octant-reduced sinl/cosl."""
    lib = _load()
    hi = np.empty(n, dtype=np.complex128)
    lo = np.empty(n, dtype=np.complex128)
    rc = lib.oracle_twiddles(n, 1 if inverse else -1, _dptr(hi.view(np.float64)),
                             _dptr(lo.view(np.float64)))
    if rc != 0:
        raise ValueError("oracle_twiddles failed")
    return DD(hi, lo)


def reference(x, inverse: bool = False) -> DD:
    """The checker used against the GPU: long-double radix-2 FFT."""
    return fft(x, inverse=inverse)


# --------------------------------------------------------------------------
# error metrics (computed in long double against a DD reference)

def rel_l2(got, ref: DD, axis=None):
    """||got - ref|| / ||ref||, per signal when axis=-1, global when None."""
    g = np.asarray(got).astype(np.clongdouble)
    r = ref.as_longdouble()
    num = np.sum(np.abs(g - r) ** 2, axis=axis)
    den = np.sum(np.abs(r) ** 2, axis=axis)
    den = np.where(den == 0, 1, den)
    return np.sqrt(num / den).astype(np.float64)


def max_abs(got, ref: DD, axis=None):
    g = np.asarray(got).astype(np.clongdouble)
    return np.max(np.abs(g - ref.as_longdouble()), axis=axis).astype(np.float64)


def mse(got, ref: DD):
    """Mean squared absolute error, the metric of src/gpu_fft.cu:41-46,62."""
    g = np.asarray(got).astype(np.clongdouble)
    return float(np.mean(np.abs(g - ref.as_longdouble()) ** 2))
