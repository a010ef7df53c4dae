"""Loads libfftoc.so (built in-tree by ``make -C fft_onchip_amd/csrc``) and
declares the C ABI of include/fftoc.h.  There is no fallback: if the library
is missing the import of the transform entry points fails loudly."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfftoc.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "fftoc.h")

_lib = None

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_u32 = ctypes.c_uint32
_int = ctypes.c_int

# name -> (argtypes, restype)
SIGNATURES = {
    "fftoc_init": ([], _int),
    "fftoc_shutdown": ([], _int),
    "fftoc_strerror": ([_int], ctypes.c_char_p),
    "fftoc_version": ([], _int),
    "fftoc_twiddle_table": ([ctypes.POINTER(_vp), ctypes.POINTER(_vp)], _int),
    "fftoc_tensor_fft": ([_vp, _vp, _i64, _int, _u32, _vp], _int),
    "fftoc_legacy8_fft": ([_vp, _vp, _i64, _int, _u32, _vp], _int),
    "fftoc_legacy16_fft": ([_vp, _vp, _i64, _int, _u32, _vp], _int),
    "fftoc_fft": ([_vp, _vp, _i64, _int, _u32, _vp], _int),
}
SIGNATURES.update({
    "fftoc_variant_count": ([_int, _u32], _int),
    "fftoc_variant_name": ([_int, _u32, _int], ctypes.c_char_p),
    "fftoc_fft_variant": ([_vp, _vp, _i64, _int, _u32, _int, _vp], _int),
    "fftoc_get_variant": ([_int, _u32, _i64], _int),
    "fftoc_set_variant": ([_int, _u32, _i64, _int], _int),
    "fftoc_autotune": ([_int, _u32, _i64, _int, _vp], _int),
    "fftoc_variant_time_us": ([_int, _u32, _int], ctypes.c_double),
    "fftoc_variant_tile": ([_int, _u32, _int], _int),
    "fftoc_variant_per_cu": ([_int, _u32, _int], _int),
    "fftoc_variant_occupancy": ([_int, _u32, _int], _int),
    "fftoc_get_workgroups_per_cu": ([_int, _u32, _i64], _int),
    "fftoc_set_workgroups_per_cu": ([_int, _u32, _i64, _int], _int),
    "fftoc_tuning_save": ([ctypes.c_char_p], _int),
    "fftoc_tuning_load": ([ctypes.c_char_p], _int),
})


class Job(ctypes.Structure):
    """struct fftoc_job (include/fftoc.h)."""
    _fields_ = [("in_", _vp), ("out", _vp), ("batch", _i64), ("n", _int), ("flags", _u32)]


SIGNATURES.update({
    "fftoc_plan_create": ([ctypes.POINTER(Job), _int, ctypes.POINTER(_vp)], _int),
    "fftoc_plan_execute": ([_vp, _vp], _int),
    "fftoc_plan_execute_graph": ([_vp, _vp], _int),
    "fftoc_plan_autotune": ([_vp, _int, _vp], _int),
    "fftoc_plan_routine": ([_vp, _int, _u32, ctypes.POINTER(ctypes.c_char_p)], _int),
    "fftoc_plan_set_routine": ([_vp, _int, _u32, ctypes.c_char_p], _int),
    "fftoc_plan_workgroups_per_cu": ([_vp], _int),
    "fftoc_plan_set_workgroups_per_cu": ([_vp, _int], _int),
    "fftoc_batch_routine_name": ([_int, _int, ctypes.POINTER(ctypes.c_char_p)], _int),
    "fftoc_plan_destroy": ([_vp], _int),
    "fftoc_batch_fft_noalloc": ([ctypes.POINTER(Job), _int, _vp], _int),
    "fftoc_server_create": ([ctypes.c_size_t, _int, ctypes.POINTER(_vp)], _int),
    "fftoc_server_submit": ([_vp, _vp, _i64, _int, _u32, ctypes.POINTER(ctypes.c_uint64)], _int),
    "fftoc_server_poll": ([_vp, ctypes.c_uint64, ctypes.POINTER(_int), ctypes.POINTER(_vp)], _int),
    "fftoc_server_release": ([_vp, ctypes.c_uint64], _int),
    "fftoc_server_gpu_time": ([_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)], _int),
    "fftoc_server_replay": ([_vp, ctypes.POINTER(Job), _int, ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)], _int),
    "fftoc_server_stats": ([_vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(_int)], _int),
    "fftoc_server_destroy": ([_vp], _int),
    "fftoc_server_set_idle_timeout": ([_vp, ctypes.c_double], _int),
})
TENSOR_SIZES = (8, 16, 64, 128, 256, 512, 4096)
for _n in TENSOR_SIZES:
    SIGNATURES[f"fftoc_tensor_fft_{_n}"] = ([_vp, _vp, _i64, _u32, _vp], _int)


class LibraryMissing(ImportError):
    pass


def load():
    """Returns the loaded ctypes library (raises LibraryMissing if not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                f"{LIB_PATH} is not built; run `make -C fft_onchip_amd/csrc` "
                "(or __graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _lib = lib
    return _lib
