"""CPU-side checks of the C ABI: the library loads, exports every symbol the
header declares, and rejects bad arguments before touching a device."""
import ctypes
import os
import re

import pytest

import fft_onchip_amd as fo
from fft_onchip_amd import _native

HEADER = _native.HEADER


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fftoc_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.LIB_PATH)
    names = declared_symbols()
    assert len(names) == len(_native.SIGNATURES)
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) <= set(_native.SIGNATURES), set(names) - set(_native.SIGNATURES)


def test_strerror_covers_every_status():
    for code, name in fo.STATUS.items():
        s = fo.lib().fftoc_strerror(code)
        assert s and s != b"unknown status", name
    assert fo.lib().fftoc_strerror(999) == b"unknown status"


def test_status_codes_match_header():
    text = open(HEADER).read()
    for name, code in re.findall(r"FFTOC_(?:ERROR_)?(\w+) = (\d+)", text):
        if name in ("SUCCESS",) or name in fo.STATUS.values():
            assert fo.STATUS[int(code)] == name


def _rc(fn, *args):
    return getattr(fo.lib(), fn)(*args)


FAKE_IN = 0x10000
FAKE_OUT = 0x20000000


@pytest.mark.parametrize("fn,n", [("fftoc_fft", 4), ("fftoc_fft", 8192), ("fftoc_fft", 100),
                                  ("fftoc_legacy8_fft", 16), ("fftoc_legacy8_fft", 128),
                                  ("fftoc_legacy16_fft", 64), ("fftoc_legacy16_fft", 512),
                                  ("fftoc_tensor_fft", 1024), ("fftoc_tensor_fft", 12)])
def test_invalid_sizes(fn, n):
    assert _rc(fn, FAKE_IN, FAKE_OUT, 1, n, 0, None) == 1


def test_argument_errors_need_no_device():
    lib = fo.lib()
    assert lib.fftoc_fft(FAKE_IN, FAKE_OUT, -1, 64, 0, None) == 2          # batch < 0
    assert lib.fftoc_fft(None, FAKE_OUT, 1, 64, 0, None) == 3              # null
    assert lib.fftoc_fft(FAKE_IN, None, 1, 64, 0, None) == 3
    assert lib.fftoc_fft(FAKE_IN, FAKE_OUT, 1, 64, 0x8, None) == 4         # unknown flag bit
    assert lib.fftoc_fft(FAKE_IN, FAKE_OUT, 1, 64, 0x30, None) == 4        # unknown precision
    assert lib.fftoc_fft(FAKE_IN, FAKE_IN + 16, 4, 64, 0, None) == 5       # partial overlap
    assert lib.fftoc_fft(FAKE_IN + 8, FAKE_OUT, 1, 64, 0, None) == 10      # misaligned
    assert lib.fftoc_legacy8_fft(FAKE_IN, FAKE_OUT, 1, 64, 0x10, None) == 6  # fp32 on legacy
    assert lib.fftoc_tensor_fft_64(None, None, 0, 0, None) == 0            # empty batch: no-op
    assert lib.fftoc_fft(None, None, 0, 4096, 0, None) == 0


def test_python_mirror_raises():
    with pytest.raises(fo.FftocError) as e:
        fo.legacy16_fft(FAKE_IN, 1, 64, out=FAKE_OUT)
    assert e.value.name == "INVALID_SIZE"
    for n in fo.TENSOR_SIZES:
        assert callable(getattr(fo, f"tensor_fft_{n}"))


def test_fused_example_library_loads_and_checks_arguments():
    """The in-LDS functor example library (include/fftoc_lds.hpp) is built and
    rejects bad arguments before any launch."""
    path = os.path.join(os.path.dirname(_native.LIB_PATH), "libfftoc_fused_example.so")
    f = ctypes.CDLL(path).fftoc_example_windowed_fft
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                  ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
    f.restype = ctypes.c_int
    assert f(None, None, 0, 64, None, 0, 1.0, 0, None) == 0      # empty batch
    assert f(None, None, -1, 64, None, 0, 1.0, 0, None) == 2     # FFTOC_ERROR_INVALID_BATCH
    assert f(None, None, 4, 64, None, 0, 1.0, 0, None) == 3      # FFTOC_ERROR_NULL_POINTER
    assert f(16, 16, 4, 64, 32, 0, 1.0, 0, None) == 5            # FFTOC_ERROR_OVERLAP
    assert f(16, 4096, 4, 48, 32, 0, 1.0, 0, None) == 1          # FFTOC_ERROR_INVALID_SIZE
    lib = ctypes.CDLL(path)
    t = lib.fftoc_example_windowed_tensor_fft
    t.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                  ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    t.restype = ctypes.c_int
    assert t(None, None, 0, 64, None, 0, 1.0, 0, 0, None) == 0
    assert t(None, None, -1, 64, None, 0, 1.0, 0, 0, None) == 2
    assert t(16, 16, 4, 64, 32, 0, 1.0, 0, 1, None) == 5
    assert t(16, 4096, 4, 32, 32, 0, 1.0, 0, 0, None) == 1      # no tensor functor at 32
    assert t(16, 4096, 4, 8, 32, 0, 1.0, 0, 1, None) == 6       # n = 8: Karatsuba only
    assert t(16, 4096, 4, 64, 32, 0, 1.0, 0, 3, None) == 4      # unknown contraction
    assert t(16, 4096, 4, 64, 32, 0, 1.0, 0, 2, None) == 6      # one matrix stage: n = 256, 4096
    o = lib.fftoc_example_onchip_fft
    o.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                  ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
    o.restype = ctypes.c_int
    assert o(None, 0, 64, None, 0, 1, 1.0, None) == 0
    assert o(None, 4, 64, None, 0, 1, 1.0, None) == 3
    assert o(16, 4, 64, 32, 4, 1, 1.0, None) == 4                # unknown family
    assert o(16, 4, 64, 32, 3, 1, 1.0, None) == 6                # one matrix stage: n = 256, 4096
    assert o(16, 4, 64, 32, 0, -1, 1.0, None) == 2               # negative repeat count
    assert o(16, 4, 100, 32, 0, 1, 1.0, None) == 1
    assert o(16, 4, 32, 32, 1, 1, 1.0, None) == 6                # no tensor functor at 32
    assert o(16, 4, 8, 32, 2, 1, 1.0, None) == 6                 # n = 8: Karatsuba only
    sh = lib.fftoc_example_onchip_fft_shape
    sh.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                   ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
    sh.restype = ctypes.c_int
    assert sh(None, 0, 64, None, 0, 1, 1, 1.0, None) == 0
    assert sh(16, 4, 64, 32, 0, 2, 1, 1.0, None) == 4            # unknown shape
    assert sh(16, 4, 64, 32, 4, 1, 1, 1.0, None) == 4            # unknown family
    assert sh(16, 4, 32, 32, 1, 1, 1, 1.0, None) == 6            # no tensor functor at 32


def test_plan_rejects_jobs_writing_into_each_other():
    """Checked before any device work: a job's output overlapping another
    job's input or output is FFTOC_ERROR_OVERLAP; shared inputs and in-place
    jobs are fine as far as validation goes."""
    J = _native.Job
    base = 1 << 20
    size = 64 * 16 * 4  # 4 signals of 64 points
    bad = [J(base, base + 8 * size, 4, 64, 0), J(base + 16 * size, base + 8 * size + 256, 4, 64, 0)]
    bad_in = [J(base, base + 8 * size, 4, 64, 0), J(base + 8 * size + 512, base + 32 * size, 4, 64, 0)]
    for jobs in (bad, bad_in):
        arr = (J * len(jobs))(*jobs)
        h = ctypes.c_void_p()
        assert fo.lib().fftoc_plan_create(arr, len(jobs), ctypes.byref(h)) == 5


def test_tile_flag_selects_a_variant_or_is_unsupported():
    """FFTOC_TILE(s) names the variant with s signals per workgroup; a tile
    the length does not offer is FFTOC_ERROR_UNSUPPORTED (checked before any
    device work), and a tile contradicting a named variant is rejected."""
    lib = fo.lib()
    n = 256
    tiles = [lib.fftoc_variant_tile(n, 0, v) for v in range(lib.fftoc_variant_count(n, 0))]
    assert all(t >= 1 for t in tiles) and lib.fftoc_variant_tile(n, 0, 999) == -1
    fake = 1 << 20  # aligned, never dereferenced: the call fails before any launch
    assert 7 not in tiles
    assert lib.fftoc_fft(fake, fake, 4, n, fo.make_flags(tile=7), None) == 6
    v = 0
    other = next(t for t in tiles if t != tiles[v])
    assert lib.fftoc_fft_variant(fake, fake, 4, n, fo.make_flags(tile=other), v, None) == 4


def test_batch_routine_table_needs_no_device():
    """The tile routines plans and servers choose from: every length has at
    least one, names are unique per length and name their length's split
    (radix factors multiplying to n), and bad arguments are rejected."""
    for n in fo.ALL_SIZES:
        names = fo.batch_routines(n)
        assert names and len(set(names)) == len(names), (n, names)
        for name in names:
            split = name.split("/")[0]
            if split.startswith("tensor:"):  # matrix-core routine: 16 x 16 x 16 = 4096
                split = split[len("tensor:"):]
            prod = 1
            for f in split.split("x"):
                prod *= int(f)
            assert prod == n, (n, name)
    lib = fo.lib()
    name = ctypes.c_char_p()
    assert lib.fftoc_batch_routine_name(100, 0, ctypes.byref(name)) == 1     # not a power of two
    assert lib.fftoc_batch_routine_name(64, 0, None) == 3                    # null output
    assert lib.fftoc_batch_routine_name(64, -1, ctypes.byref(name)) == 13    # no such index
    assert lib.fftoc_plan_set_routine(None, 64, 0, b"8x8/tile/S32") == 13    # no plan


def test_variant_lists_per_data_type_need_no_device(tmp_path):
    """Each length has a complex128 and a complex64 variant list; the
    enumeration calls select one by FFTOC_DATA_C64, names are unique within a
    list, every tile is positive, and the persisted tuning table keys its lines
    by data type (written and read back without a device)."""
    lib = fo.lib()
    for n in fo.ALL_SIZES:
        for c64 in (False, True):
            dt = fo.DATA_C64 if c64 else 0
            names = fo.variants(n, c64=c64)
            assert names and len(set(names)) == len(names), (n, c64)
            assert lib.fftoc_variant_name(n, dt, len(names)) is None
            assert all(lib.fftoc_variant_tile(n, dt, v) >= 1 for v in range(len(names)))
            assert 0 <= lib.fftoc_get_variant(n, dt, 1 << 12) < len(names)
            assert fo.variant_index(n, names[-1], c64=c64) == len(names) - 1
        # complex64 lists hold no tensor variants (FP64 MFMA only)
        assert not any(v.startswith("tensor:") for v in fo.variants(n, c64=True))
    assert lib.fftoc_variant_count(100, 0) == 0
    path = str(tmp_path / "tune.txt").encode()
    assert lib.fftoc_tuning_save(path) == 0
    lines = [ln.split() for ln in open(path) if not ln.startswith("#")]
    assert {ln[2] for ln in lines} == {"c128", "c64"}
    assert lib.fftoc_tuning_load(path) == 0
    bad = tmp_path / "bad.txt"
    bad.write_text("fft 64 c32 3 radix:8x8/tile/S16\nfft 64 3 radix:8x8/tile/S16\n")
    assert lib.fftoc_tuning_load(str(bad).encode()) == 13


def test_workgroups_per_cu_setting_and_tuning_lines_need_no_device(tmp_path):
    """The workgroups-per-CU cap of fftoc_fft's choice (the reference's
    blocks-per-SM sweep, src/saturation.cu:15-62) is kept per batch bucket,
    reset by fftoc_set_variant, range-checked, and persisted as a `wgs=<k>`
    token that fftoc_tuning_load reads back; a malformed token is rejected."""
    lib = fo.lib()
    n, dt = 256, fo.DATA_C64
    assert lib.fftoc_get_workgroups_per_cu(n, dt, 1 << 10) == 0
    assert lib.fftoc_set_workgroups_per_cu(n, dt, 1 << 10, 64) == 4       # above 63
    assert lib.fftoc_set_workgroups_per_cu(n, dt, 1 << 10, -1) == 4
    assert lib.fftoc_set_workgroups_per_cu(100, dt, 1 << 10, 2) == 1      # bad length
    assert lib.fftoc_set_workgroups_per_cu(n, dt, 1 << 10, 3) == 0
    assert lib.fftoc_get_workgroups_per_cu(n, dt, 1 << 10) == 3
    assert lib.fftoc_get_workgroups_per_cu(n, dt, 1 << 11) == 0           # other bucket
    assert lib.fftoc_get_workgroups_per_cu(n, 0, 1 << 10) == 0            # other data type
    path = str(tmp_path / "tune.txt").encode()
    assert lib.fftoc_tuning_save(path) == 0
    want = f"fft {n} c64 10 wgs=3 "
    assert any(ln.startswith(want) for ln in open(path)), "cap not persisted"
    v = lib.fftoc_get_variant(n, dt, 1 << 10)
    assert lib.fftoc_set_variant(n, dt, 1 << 10, v) == 0
    assert lib.fftoc_get_workgroups_per_cu(n, dt, 1 << 10) == 0            # reset with the variant
    assert lib.fftoc_tuning_load(path) == 0
    assert lib.fftoc_get_workgroups_per_cu(n, dt, 1 << 10) == 3            # restored
    assert lib.fftoc_get_variant(n, dt, 1 << 10) == v
    name = fo.variants(n, c64=True)[v]
    bad = tmp_path / "bad.txt"
    bad.write_text(f"fft {n} c64 10 wgs=99 {name}\nfft {n} c64 10 wgs= {name}\n")
    assert lib.fftoc_tuning_load(str(bad).encode()) == 13
    assert lib.fftoc_get_workgroups_per_cu(n, dt, 1 << 10) == 3            # bad lines skipped
    assert lib.fftoc_set_workgroups_per_cu(n, dt, -1, 0) == 0
    assert lib.fftoc_plan_workgroups_per_cu(None) == -1
    assert lib.fftoc_plan_set_workgroups_per_cu(None, 2) == 13
    assert fo.make_flags(per_cu=5) == 5 << 24
    fake = 1 << 20  # aligned, never dereferenced: rejected before any launch
    assert lib.fftoc_fft(fake, fake, 4, 100, fo.make_flags(per_cu=5), None) == 1
