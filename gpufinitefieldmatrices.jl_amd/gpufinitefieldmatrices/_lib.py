"""ctypes binding of libmodp_hip.so (the C ABI declared in include/modp.h).

The library must be built (``python gpufinitefieldmatrices.jl_amd/build.py``);
loading fails loudly otherwise -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmodp_hip.so")

_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_int = ctypes.c_int
_vp = ctypes.c_void_p
_p64 = ctypes.POINTER(ctypes.c_int64)
_sz = ctypes.c_size_t

MODP_OK = 0
MODP_ERR_DIM = 1
MODP_ERR_MODULUS_RANGE = 2
MODP_ERR_MODULUS_MISMATCH = 3
MODP_ERR_NOT_INVERTIBLE = 4
MODP_ERR_DTYPE = 5
MODP_ERR_OOM = 6
MODP_ERR_DEVICE = 7
MODP_ERR_ARG = 8
MODP_ERR_NO_DEVICE = 9

MODP_F64 = 0
MODP_F32 = 1

FORM_AUTO = 0
FORM_PLAIN_F64 = 1
FORM_LIMB_I8 = 2
FORM_RNS_I8 = 3

TUNE_OFF = 0
TUNE_CACHED = 1
TUNE_ON_MISS = 2

ELIM_PLUQ = 0
ELIM_LU = 1
ELIM_RREF = 2


class GemmJob(ctypes.Structure):
    _fields_ = [("m", _i64), ("n", _i64), ("k", _i64), ("p", _i64),
                ("A", _vp), ("lda", _i64), ("B", _vp), ("ldb", _i64),
                ("C", _vp), ("ldc", _i64)]


class GemmConfig(ctypes.Structure):
    _fields_ = [("form", ctypes.c_int32), ("tile_m", ctypes.c_int32), ("tile_n", ctypes.c_int32),
                ("ksplit", ctypes.c_int32)]


class ElimJob(ctypes.Structure):
    _fields_ = [("m", _i64), ("n", _i64), ("p", _i64),
                ("A", _vp), ("lda", _i64), ("U", _vp), ("ldu", _i64),
                ("L", _vp), ("ldl", _i64),
                ("row_info", _vp), ("col_info", _vp), ("counts", _vp)]


# (name, restype, argtypes) of every exported symbol of include/modp.h
SIGNATURES = [
    ("modp_version", ctypes.c_char_p, []),
    ("modp_device_count", _int, [ctypes.POINTER(_int)]),
    ("modp_ctx_create", _int, [_int, ctypes.POINTER(_vp)]),
    ("modp_ctx_destroy", _int, [_vp]),
    ("modp_last_error", ctypes.c_char_p, [_vp]),
    ("modp_synchronize", _int, [_vp]),
    ("modp_stream", _vp, [_vp]),
    ("modp_malloc", _int, [_vp, _sz, ctypes.POINTER(_vp)]),
    ("modp_free", _int, [_vp, _vp]),
    ("modp_memcpy_h2d", _int, [_vp, _vp, _vp, _sz]),
    ("modp_memcpy_d2h", _int, [_vp, _vp, _vp, _sz]),
    ("modp_memcpy_d2d", _int, [_vp, _vp, _vp, _sz]),
    ("modp_memset_zero", _int, [_vp, _vp, _sz]),
    ("modp_mod", _int, [_vp, _int, _i64, _i64, _vp, _i64, _i64]),
    ("modp_rand", _int, [_vp, _int, _i64, _i64, _vp, _i64, _i64, _u64, _u64]),
    ("modp_eye", _int, [_vp, _int, _i64, _i64, _vp, _i64]),
    ("modp_gemm", _int, [_vp, _int, _int, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i64]),
    ("modp_gemm_noalloc", _int, [_vp, _int, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i64]),
    ("modp_gemv", _int, [_vp, _int, _i64, _i64, _vp, _i64, _vp, _vp, _i64]),
    ("modp_autotune_mode", _int, [_vp, _int]),
    ("modp_autotune_gemm", _int, [_vp, _int, _i64, _i64, _i64, _i64, ctypes.POINTER(GemmConfig),
                                  ctypes.POINTER(ctypes.c_double)]),
    ("modp_gemm_config_get", _int, [_vp, _int, _i64, _i64, _i64, _i64, ctypes.POINTER(GemmConfig),
                                    ctypes.POINTER(_int)]),
    ("modp_gemm_ex", _int, [_vp, ctypes.POINTER(GemmConfig), _int, _i64, _i64, _i64, _vp, _i64, _vp, _i64,
                            _vp, _i64, _i64]),
    ("modp_autotune_export", _int, [_vp, ctypes.c_char_p, _sz, ctypes.POINTER(_sz)]),
    ("modp_autotune_import", _int, [_vp, ctypes.c_char_p]),
    ("modp_pluq", _int, [_vp, _int, _int, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i64,
                         _p64, _p64, _p64, _p64, _p64]),
    ("modp_lu", _int, [_vp, _int, _int, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i64,
                       _p64, _p64, _p64]),
    ("modp_rref", _int, [_vp, _int, _int, _i64, _i64, _vp, _i64, _vp, _i64, _i64, _p64, _p64]),
    ("modp_crt", _int, [_vp, _int, _i64, _i64, _int, ctypes.POINTER(_vp), _p64, _p64, _int, _vp, _i64, _i64]),
    ("modp_crt2", _int, [_vp, _int, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64]),
    ("modp_trtri", _int, [_vp, _int, _int, _i64, _i64, _vp, _i64, _vp, _i64, _i64]),
    ("modp_inverse", _int, [_vp, _int, _int, _i64, _vp, _i64, _vp, _i64, _i64, ctypes.POINTER(_int)]),
    ("modp_apply_row_perm", _int, [_vp, _int, _i64, _i64, _vp, _i64, _p64, _i64, _int]),
    ("modp_apply_col_perm", _int, [_vp, _int, _i64, _i64, _vp, _i64, _p64, _i64, _int]),
    ("modp_gemm_batched_plan", _int, [_vp, _int, ctypes.POINTER(GemmJob), _i64, ctypes.POINTER(_vp)]),
    ("modp_gemm_batched_run", _int, [_vp, _vp]),
    ("modp_gemm_batched_destroy", _int, [_vp, _vp]),
    ("modp_gemm_batched", _int, [_vp, _int, ctypes.POINTER(GemmJob), _i64]),
    ("modp_elim_batched", _int, [_vp, _int, _int, ctypes.POINTER(ElimJob), _i64]),
    ("modp_elim_batched_plan", _int, [_vp, _int, _int, ctypes.POINTER(ElimJob), _i64, ctypes.POINTER(_vp)]),
    ("modp_elim_batched_run", _int, [_vp, _vp]),
    ("modp_elim_batched_destroy", _int, [_vp, _vp]),
    ("modp_kernel_timing", _int, [_vp, _int]),
    ("modp_kernel_time", _int, [_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), _p64]),
    ("modp_kernel_timing_reset", _int, [_vp]),
    ("modp_event_record", _int, [_vp, _int]),
    ("modp_event_elapsed", _int, [_vp, _int, _int, ctypes.POINTER(ctypes.c_float)]),
    ("modp_probe_peak", _int, [_vp, _int, ctypes.POINTER(ctypes.c_double)]),
]

_LIB = None


def load():
    """Load libmodp_hip.so and declare every entry point's signature."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python gpufinitefieldmatrices.jl_amd/build.py` "
                "(the HIP library is the only implementation: there is no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _LIB = lib
    return _LIB
