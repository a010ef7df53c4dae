"""GPU parity: every path and size against the long-double oracle.

Tolerance (FP64): relative L2 error per signal <= 1e-13, this path's
accuracy target; observed errors are around 1e-16.
"""
import numpy as np
import pytest

import fft_onchip_amd as fo
import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-13

PATH_SIZES = ([("tensor", n) for n in fo.TENSOR_SIZES + (32,)]
              + [("legacy8", n) for n in fo.LEGACY8_SIZES]
              + [("legacy16", n) for n in fo.LEGACY16_SIZES]
              + [("best", n) for n in fo.ALL_SIZES])


def signals(n, batch, seed, kind="uniform"):
    g = np.random.Generator(np.random.Philox(seed))
    if kind == "uniform":
        return g.uniform(-1, 1, (batch, n)) + 1j * g.uniform(-1, 1, (batch, n))
    return g.standard_normal((batch, n)) + 1j * g.standard_normal((batch, n))


def batch_for(n):
    # not a multiple of any tile size, small enough for the oracle
    return max(3, min(515, 65536 // n)) | 1


@pytest.fixture(scope="module", autouse=True)
def _device():
    assert fo.device_count() >= 1, "no HIP device visible"
    fo.init()


@pytest.mark.parametrize("path,n", PATH_SIZES)
def test_forward_matches_oracle(path, n):
    x = signals(n, batch_for(n), seed=n)
    print(f"[parity] path={path} n={n} batch={x.shape[0]}", flush=True)
    got = fo.fft_host(x, path=path)
    err = oracle.rel_l2(got, oracle.reference(x), axis=-1)
    assert np.max(err) <= TOL, (path, n, np.max(err))


# Defaults that take over only from a batch size on (one-wave lane
# workgroups; the staged and the persistent one-wave matrix-core kernels with
# their workgroup caps), each at its first batch bucket plus 3 signals (a
# partial last tile): forward against the oracle on the first and last 32
# signals, then the normalised inverse of that output gives every input back.
LARGE_BATCH_DEFAULTS = [("best", 8, 1 << 18, "radix:8/lanes/S8/64 threads"),
                        ("best", 16, 1 << 17, "radix:16/lanes/S4/64 threads"),
                        ("best", 32, 1 << 16, "radix:32/lanes/S4/64 threads"),
                        ("tensor", 8, 1 << 18, None), ("tensor", 128, 1 << 15, None),
                        ("tensor", 256, 1 << 14, None),
                        ("tensor", 512, 1 << 14, None)]


@pytest.mark.parametrize("path,n,batch,variant", LARGE_BATCH_DEFAULTS)
def test_large_batch_defaults_forward_and_normalized_inverse(path, n, batch, variant):
    x = signals(n, batch + 3, seed=1700 + n)
    print(f"[large-batch default] path={path} n={n} batch={x.shape[0]}", flush=True)
    if variant is not None:
        assert fo.variants(n)[fo.lib().fftoc_get_variant(n, 0, x.shape[0])] == variant
    fn = fo.fft if path == "best" else fo.tensor_fft
    src = fo.DeviceBuffer.from_array(x)
    dst = fo.DeviceBuffer(x.nbytes)
    fn(src, x.shape[0], n, out=dst)
    fo.synchronize()
    got = dst.download(x.shape)
    idx = np.r_[0:32, x.shape[0] - 32:x.shape[0]]
    err = np.max(oracle.rel_l2(got[idx], oracle.reference(x[idx]), axis=-1))
    assert err <= TOL, (path, n, err)
    fn(dst, x.shape[0], n, out=src, flags=fo.INVERSE | fo.NORMALIZE)
    fo.synchronize()
    assert np.max(np.abs(src.download(x.shape) - x)) < 1e-13, (path, n)


@pytest.mark.parametrize("path,n", PATH_SIZES)
def test_inverse_and_normalize(path, n):
    x = signals(n, 33, seed=100 + n)
    ref = oracle.reference(x, inverse=True)
    got = fo.fft_host(x, path=path, inverse=True)
    assert np.max(oracle.rel_l2(got, ref, axis=-1)) <= TOL
    gotn = fo.fft_host(x, path=path, inverse=True, normalize=True)
    refn = oracle.DD(ref.hi / n, ref.lo / n)
    assert np.max(oracle.rel_l2(gotn, refn, axis=-1)) <= TOL
    # round trip: forward then normalised inverse gives x back
    back = fo.fft_host(fo.fft_host(x, path=path), path=path, inverse=True, normalize=True)
    assert np.max(np.abs(back - x)) < 1e-13


@pytest.mark.parametrize("path,n", [("tensor", 64), ("tensor", 4096), ("legacy16", 256),
                                    ("legacy8", 512), ("best", 2048), ("best", 8), ("best", 4096),
                                    ("best", 64)])
def test_in_place(path, n):
    # best 4096 / 64 run persistent kernels whose workgroups read their next
    # signals before storing the current ones: batches of several grids
    batch = {("best", 4096): 1031, ("best", 64): 70001}.get((path, n), 77)
    x = signals(n, batch, seed=7)
    got = fo.fft_host(x, path=path, in_place=True)
    assert np.max(oracle.rel_l2(got, oracle.reference(x), axis=-1)) <= TOL


@pytest.mark.parametrize("path,n", PATH_SIZES)
def test_golden_fixtures(path, n, golden):
    if n not in golden["sizes"]:
        pytest.skip("no fixture for this size")
    x = golden[f"x_{n}"]
    ref = oracle.DD(golden[f"fwd_hi_{n}"], golden[f"fwd_lo_{n}"])
    got = fo.fft_host(x, path=path)
    assert np.max(oracle.rel_l2(got, ref, axis=-1)) <= TOL


@pytest.mark.parametrize("path", ["tensor", "best", "legacy16", "legacy8"])
def test_batch_invariance(path):
    """Each signal's result is bit-identical alone and inside a batch."""
    for n in {"tensor": (8, 64, 128, 4096), "best": (16, 256, 1024), "legacy16": (16, 4096),
              "legacy8": (8, 512)}[path]:
        x = signals(n, 257, seed=3)
        full = fo.fft_host(x, path=path)
        for i in (0, 1, 130, 256):
            alone = fo.fft_host(x[i:i + 1], path=path)
            assert np.array_equal(alone[0], full[i]), (path, n, i)


@pytest.mark.parametrize("path", ["tensor", "best", "legacy16", "legacy8"])
def test_pure_tones_land_in_their_bins(path):
    n = 4096
    rng = np.random.default_rng(7)
    k0 = rng.integers(0, n, 64)
    j = np.arange(n)
    x = np.exp(2j * np.pi * ((k0[:, None] * j[None, :]) % n) / n)
    X = fo.fft_host(x, path=path)
    assert np.array_equal(np.argmax(np.abs(X), axis=-1), k0)
    off = X.copy()
    off[np.arange(64), k0] = 0
    assert np.max(np.abs(off)) < 1e-9
    assert np.max(np.abs(X[np.arange(64), k0] - n)) < 1e-9


def _variants(n):
    return list(enumerate(fo.variants(n)))


@pytest.mark.parametrize("n", fo.ALL_SIZES)
def test_every_variant_matches_oracle(n):
    """Every kernel variant the tuner can pick, forward and inverse, including
    a batch that leaves a partial last tile."""
    x = signals(n, batch_for(n), seed=500 + n)
    ref_f = oracle.reference(x)
    ref_i = oracle.reference(x, inverse=True)
    src = fo.DeviceBuffer.from_array(x)
    dst = fo.DeviceBuffer(x.nbytes)
    for v, name in _variants(n):
        for inv, ref in ((False, ref_f), (True, ref_i)):
            print(f"[variant] n={n} v={v} {name} inverse={inv}", flush=True)
            rc = fo.lib().fftoc_fft_variant(src.ptr, dst.ptr, x.shape[0], n, fo.INVERSE if inv else 0,
                                            v, None)
            assert rc == 0, (name, rc)
            fo.synchronize()
            got = dst.download(x.shape)
            err = np.max(oracle.rel_l2(got, ref, axis=-1))
            assert err <= TOL, (n, name, inv, err)


@pytest.mark.parametrize("n", fo.ALL_SIZES)
def test_persistent_variants_over_several_grids(n):
    """Persistent and pipelined variants at a batch several times what their
    grid holds at once, so every workgroup loops and prefetches; out of place
    and in place (the prefetch must read each next signal before any of it is
    overwritten)."""
    loops = [(v, name) for v, name in _variants(n) if "persistent" in name or "pipelined" in name]
    if not loops:
        pytest.skip(f"no persistent variant at n={n}")
    x = signals(n, (1 << 23) // n + 7, seed=900 + n)
    ref_f = oracle.reference(x)
    ref_i = oracle.reference(x, inverse=True)
    src = fo.DeviceBuffer.from_array(x)
    dst = fo.DeviceBuffer(x.nbytes)
    for v, name in loops:
        for inv, ref in ((False, ref_f), (True, ref_i)):
            for in_place in (False, True):
                print(f"[persistent] n={n} v={v} {name} inverse={inv} in_place={in_place}", flush=True)
                src.upload(x)  # an earlier in-place call overwrote it
                out = src if in_place else dst
                rc = fo.lib().fftoc_fft_variant(src.ptr, out.ptr, x.shape[0], n,
                                                fo.INVERSE if inv else 0, v, None)
                assert rc == 0, (name, rc)
                fo.synchronize()
                err = np.max(oracle.rel_l2(out.download(x.shape), ref, axis=-1))
                assert err <= TOL, (n, name, inv, in_place, err)


@pytest.mark.parametrize("n", [64, 256, 4096])
def test_tile_flag_runs_a_variant_with_that_tile(n):
    """fftoc_fft / fftoc_tensor_fft with FFTOC_TILE(s) run a variant with s
    signals per workgroup: every tile the length offers matches the oracle,
    and the tile of fftoc_fft's own choice reproduces its output exactly."""
    lib = fo.lib()
    x = signals(n, batch_for(n), seed=900 + n)
    src = fo.DeviceBuffer.from_array(x)
    dst = fo.DeviceBuffer(x.nbytes)
    ref = oracle.reference(x)
    for fam, fn in (("radix:", lib.fftoc_fft), ("tensor:", lib.fftoc_tensor_fft)):
        tiles = sorted({lib.fftoc_variant_tile(n, 0, v) for v, name in _variants(n)
                        if name.startswith(fam) and lib.fftoc_variant_tile(n, 0, v) <= 255})
        for tile in tiles:
            print(f"[tile] n={n} {fam} tile={tile}", flush=True)
            assert fn(src.ptr, dst.ptr, x.shape[0], n, fo.make_flags(tile=tile), None) == 0
            fo.synchronize()
            assert np.max(oracle.rel_l2(dst.download(x.shape), ref, axis=-1)) <= TOL, (fam, tile)
    chosen = lib.fftoc_get_variant(n, 0, x.shape[0])
    tile = lib.fftoc_variant_tile(n, 0, chosen)
    plain = fo.fft_host(x)
    if tile <= 255:
        assert lib.fftoc_fft(src.ptr, dst.ptr, x.shape[0], n, fo.make_flags(tile=tile), None) == 0
        fo.synchronize()
        assert np.array_equal(dst.download(x.shape), plain)


@pytest.mark.parametrize("c64", [False, True])
def test_autotune_and_tuning_roundtrip(tmp_path, c64):
    """fftoc_autotune times every variant of the data type's list and makes
    the fastest the default of its batch bucket; the persisted table restores
    that choice for that data type only."""
    lib = fo.lib()
    n = 64 if c64 else 256  # lengths with several variants of the data type
    dt = fo.DATA_C64 if c64 else 0
    count = lib.fftoc_variant_count(n, dt)
    assert count == len(fo.variants(n, c64=c64)) >= 2
    other_dt = 0 if c64 else fo.DATA_C64
    before_other = lib.fftoc_get_variant(n, other_dt, 4096)
    assert lib.fftoc_autotune(n, dt, 4096, 3, None) == 0
    times = [lib.fftoc_variant_time_us(n, dt, v) for v in range(count)]
    assert all(t > 0 for t in times)
    best = lib.fftoc_get_variant(n, dt, 4096)
    assert times[best] == min(times)
    assert lib.fftoc_get_variant(n, other_dt, 4096) == before_other  # other list untouched
    path = str(tmp_path / "tune.txt").encode()
    assert lib.fftoc_tuning_save(path) == 0
    text = open(path).read()
    cap = lib.fftoc_get_workgroups_per_cu(n, dt, 4096)
    assert cap == lib.fftoc_variant_per_cu(n, dt, best)
    assert 0 <= cap < max(2, lib.fftoc_variant_occupancy(n, dt, best))
    wgs = f"wgs={cap} " if cap else ""
    assert f"fft {n} {'c64' if c64 else 'c128'} 12 {wgs}{fo.variants(n, c64=c64)[best]}\n" in text
    other = (best + 1) % count
    assert lib.fftoc_set_variant(n, dt, 4096, other) == 0
    assert lib.fftoc_get_variant(n, dt, 4096) == other
    assert lib.fftoc_tuning_load(path) == 0
    assert lib.fftoc_get_variant(n, dt, 4096) == best
    x = signals(n, 64, seed=9)
    if c64:
        x = x.astype(np.complex64)
    tol = 2e-6 if c64 else TOL
    got = fo.fft_host(x)
    assert np.max(oracle.rel_l2(got, oracle.reference(x.astype(np.complex128)), axis=-1)) <= tol


@pytest.mark.parametrize("n,c64", [(8, False), (64, False), (128, False), (256, False), (512, False),
                                   (4096, False),
                                   (64, True), (128, True), (256, True), (512, True),
                                   (4096, True)])
def test_workgroups_per_cu_caps_keep_results_bit_for_bit(n, c64):
    """FFTOC_WGS_PER_CU(k) holds a variant to k workgroups per CU -- a
    persistent grid of k per CU, or (one workgroup per tile) an LDS request
    that lets only k fit -- and never changes a result: every variant at cap
    1 and at one below its occupancy reproduces its uncapped output bit for
    bit, on a batch several times one grid."""
    lib = fo.lib()
    dt = fo.DATA_C64 if c64 else 0
    x = signals(n, (1 << 21) // n + 7, seed=1300 + n)
    if c64:
        x = x.astype(np.complex64)
    src = fo.DeviceBuffer.from_array(x)
    dst = fo.DeviceBuffer(x.nbytes)
    names = fo.variants(n, c64=c64)
    for v, name in enumerate(names):
        occ = lib.fftoc_variant_occupancy(n, dt, v)
        assert 1 <= occ <= 63, (name, occ)
        assert lib.fftoc_fft_variant(src.ptr, dst.ptr, x.shape[0], n, dt, v, None) == 0
        fo.synchronize()
        plain = dst.download(x.shape, x.dtype)
        for cap in sorted({1, max(1, occ - 1)}):
            print(f"[per-cu] n={n} c64={c64} v={v} {name} occupancy={occ} cap={cap}", flush=True)
            dst.upload(np.zeros_like(x))
            rc = lib.fftoc_fft_variant(src.ptr, dst.ptr, x.shape[0], n,
                                       dt | fo.make_flags(per_cu=cap), v, None)
            assert rc == 0, (name, cap, rc)
            fo.synchronize()
            assert np.array_equal(dst.download(x.shape, x.dtype), plain), (name, cap)
    tol = 2e-6 if c64 else TOL
    err = np.max(oracle.rel_l2(plain[:64], oracle.reference(x[:64].astype(np.complex128)), axis=-1))
    assert err <= tol


@pytest.mark.parametrize("n", fo.TENSOR_SIZES)
def test_per_size_tensor_aliases_on_data(n):
    """Each per-size entry point fftoc_tensor_fft_N (the drop-in for
    tensor_fft_N<..., N, ...>) transforms data like fftoc_tensor_fft,
    forward and normalised inverse, out of place and in place."""
    x = signals(n, batch_for(n), seed=300 + n)
    src = fo.DeviceBuffer.from_array(x)
    dst = fo.DeviceBuffer(x.nbytes)
    alias = getattr(fo, f"tensor_fft_{n}")
    print(f"[alias] fftoc_tensor_fft_{n} batch={x.shape[0]}", flush=True)
    alias(src, x.shape[0], out=dst)
    fo.synchronize()
    got = dst.download(x.shape)
    assert np.max(oracle.rel_l2(got, oracle.reference(x), axis=-1)) <= TOL, n
    generic = fo.fft_host(x, path="tensor")
    assert np.array_equal(got, generic), n
    alias(src, x.shape[0], flags=fo.INVERSE | fo.NORMALIZE)   # in place
    fo.synchronize()
    ref = oracle.reference(x, inverse=True)
    got = src.download(x.shape)
    assert np.max(oracle.rel_l2(got, oracle.DD(ref.hi / n, ref.lo / n), axis=-1)) <= TOL, n
