"""GPU parity of the device-callable in-LDS functors (include/fftoc_lds.hpp)
through the fused example library: Hann window applied while loading into
LDS, fftoc::LdsFft in place, store.  Checked against the oracle applied to
the windowed signals, FP64 tolerance 1e-13 relative L2 per signal."""
import ctypes
import os

import numpy as np
import pytest

import fft_onchip_amd as fo
import oracle
from fft_onchip_amd.hip import download_ptr

pytestmark = pytest.mark.gpu
TOL = 1e-13
EXAMPLE = os.path.join(os.path.dirname(fo.__file__), "libfftoc_fused_example.so")

# signals per workgroup of the example kernel (see fused_example.hip)
PER_GROUP = {8: 128, 16: 64, 32: 64, 64: 32, 128: 32, 256: 16, 512: 4, 1024: 4, 2048: 2, 4096: 1}


@pytest.fixture(scope="module")
def example():
    assert fo.device_count() >= 1, "no HIP device visible"
    fo.init()
    lib = ctypes.CDLL(EXAMPLE)
    f = lib.fftoc_example_windowed_fft
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                  ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
    f.restype = ctypes.c_int
    tw = ctypes.c_void_p()
    assert fo.lib().fftoc_twiddle_table(ctypes.byref(tw), None) == 0 and tw.value
    return f, tw


def signals(n, batch, seed):
    g = np.random.Generator(np.random.Philox(seed))
    return g.uniform(-1, 1, (batch, n)) + 1j * g.uniform(-1, 1, (batch, n))


@pytest.mark.parametrize("n", fo.ALL_SIZES)
@pytest.mark.parametrize("window,inverse", [(False, False), (True, False), (True, True)])
def test_in_lds_functor_matches_oracle(example, n, window, inverse):
    f, tw = example
    batch = 2 * PER_GROUP[n] + 1  # a partial last workgroup
    x = signals(n, batch, seed=700 + n)
    xw = x * (0.5 - 0.5 * np.cos(2 * np.pi * np.arange(n) / n)) if window else x
    scale = 1.0 / n if inverse else 1.0
    print(f"[lds] n={n} batch={batch} window={window} inverse={inverse}", flush=True)
    src = fo.DeviceBuffer.from_array(x)
    dst = fo.DeviceBuffer(x.nbytes)
    assert f(src.ptr, dst.ptr, batch, n, tw, int(inverse), scale, int(window), None) == 0
    fo.synchronize()
    got = dst.download(x.shape)
    ref = oracle.reference(xw, inverse=inverse)
    if inverse:
        ref = oracle.DD(ref.hi / n, ref.lo / n)
    err = np.max(oracle.rel_l2(got, ref, axis=-1))
    assert err <= TOL, (n, window, inverse, err)


def test_twiddle_table_entry_point():
    assert fo.device_count() >= 1
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert fo.lib().fftoc_twiddle_table(ctypes.byref(a), ctypes.byref(b)) == 0
    assert a.value and b.value and a.value != b.value
    t = download_ptr(a.value, (4096,), np.complex128)
    assert np.max(np.abs(t - np.exp(-2j * np.pi * np.arange(4096) / 4096))) < 1e-15
    assert fo.lib().fftoc_twiddle_table(None, None) == 3  # FFTOC_ERROR_NULL_POINTER
