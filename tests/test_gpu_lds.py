"""GPU parity of the device-callable in-LDS functors (include/fftoc_lds.hpp)
through the fused example library: Hann window applied while loading into
LDS, fftoc::LdsFft (radix passes) or fftoc::LdsTensorFft (matrix cores, both
contractions) in place, store; and the on-chip timing harness (one load, R
transforms in LDS, one store).  Checked against the oracle applied to the
windowed signals, FP64 tolerance 1e-13 relative L2 per signal."""
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
    t = lib.fftoc_example_windowed_tensor_fft
    t.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                  ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    t.restype = ctypes.c_int
    o = lib.fftoc_example_onchip_fft
    o.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                  ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
    o.restype = ctypes.c_int
    sh = lib.fftoc_example_onchip_fft_shape
    sh.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                   ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
    sh.restype = ctypes.c_int
    f.tensor, f.onchip, f.onchip_shape = t, o, sh
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


# signals per workgroup of the example's tensor functors (fused_example.hip)
TENSOR_PER_GROUP = {8: 256, 16: 64, 64: 32, 128: 16, 256: 8, 512: 4, 4096: 1}
# contraction: 0 Karatsuba, 1 two-product (not at n = 8), 2 Karatsuba on one
# matrix stage of the 256-point core (n = 256, 4096)
TENSOR_CASES = [(n, c) for n in sorted(TENSOR_PER_GROUP) for c in (0, 1, 2)
                if not (n == 8 and c == 1) and not (c == 2 and n not in (256, 4096))]


@pytest.mark.parametrize("n,two_product", TENSOR_CASES)
@pytest.mark.parametrize("window,inverse", [(False, False), (True, True)])
def test_in_lds_tensor_functor_matches_oracle(example, n, two_product, window, inverse):
    """fftoc::LdsTensorFft<N, UNITS, FPU, CONTRACTION> (the reference's
    tensor_fft_N functor form on the matrix cores), several units per
    workgroup, a partial last workgroup."""
    f, tw = example
    batch = 2 * TENSOR_PER_GROUP[n] + 1
    x = signals(n, batch, seed=800 + n)
    xw = x * (0.5 - 0.5 * np.cos(2 * np.pi * np.arange(n) / n)) if window else x
    scale = 1.0 / n if inverse else 1.0
    print(f"[lds tensor] n={n} two_product={two_product} window={window} inverse={inverse}", flush=True)
    src = fo.DeviceBuffer.from_array(x)
    dst = fo.DeviceBuffer(x.nbytes)
    assert f.tensor(src.ptr, dst.ptr, batch, n, tw, int(inverse), scale, int(window),
                    int(two_product), None) == 0
    fo.synchronize()
    got = dst.download(x.shape)
    ref = oracle.reference(xw, inverse=inverse)
    if inverse:
        ref = oracle.DD(ref.hi / n, ref.lo / n)
    err = np.max(oracle.rel_l2(got, ref, axis=-1))
    assert err <= TOL, (n, two_product, window, inverse, err)


@pytest.mark.parametrize("n,family", [(8, 0), (64, 0), (128, 0), (256, 0), (512, 0), (4096, 0),
                                      (8, 1), (16, 1), (64, 1), (128, 1), (256, 1), (512, 1),
                                      (4096, 1), (64, 2), (128, 2), (512, 2), (4096, 2),
                                      (256, 3), (4096, 3)])
def test_onchip_harness_repeats_in_lds(example, n, family):
    """The on-chip timing harness: R = 1 and 2 transforms of the LDS-resident
    signals, each output times 1/sqrt(N), equal the oracle applied R times."""
    f, tw = example
    batch = 67
    x = signals(n, batch, seed=900 + n)
    scale = 1.0 / np.sqrt(n)
    ref = x
    for repeats in (1, 2):
        print(f"[onchip] n={n} family={family} repeats={repeats}", flush=True)
        buf = fo.DeviceBuffer.from_array(x)
        assert f.onchip(buf.ptr, batch, n, tw, family, repeats, scale, None) == 0
        fo.synchronize()
        got = buf.download(x.shape)
        r = oracle.reference(ref)
        ref = r.as_longdouble().astype(np.complex128) * scale  # input of the next repeat
        err = np.max(oracle.rel_l2(got, oracle.DD(r.hi * scale, r.lo * scale), axis=-1))
        assert err <= TOL, (n, family, repeats, err)


@pytest.mark.parametrize("n,family", [(8, 0), (64, 0), (256, 0), (4096, 0), (8, 1), (16, 1),
                                      (64, 1), (128, 1), (256, 1), (512, 1), (4096, 1), (16, 2),
                                      (64, 2), (128, 2), (256, 2), (512, 2), (256, 3)])
def test_onchip_harness_one_wave_shape(example, n, family):
    """The harness's one-wave workgroup shape (the fewest signals a functor
    unit takes; used at a few FFTs per CU) transforms like the oracle, over a
    batch that leaves a partial last workgroup."""
    f, tw = example
    batch = 37
    x = signals(n, batch, seed=950 + n)
    scale = 1.0 / np.sqrt(n)
    print(f"[onchip shape 1] n={n} family={family}", flush=True)
    buf = fo.DeviceBuffer.from_array(x)
    assert f.onchip_shape(buf.ptr, batch, n, tw, family, 1, 1, scale, None) == 0
    fo.synchronize()
    r = oracle.reference(x)
    err = np.max(oracle.rel_l2(buf.download(x.shape), oracle.DD(r.hi * scale, r.lo * scale), axis=-1))
    assert err <= TOL, (n, family, err)
