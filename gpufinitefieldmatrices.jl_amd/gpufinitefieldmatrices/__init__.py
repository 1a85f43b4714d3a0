"""Host mirror of GPUFiniteFieldMatrices.jl's CuModMatrix hot path on MI355X.

Names, argument meaning and error behaviour follow the reference package
(src/GPUFiniteFieldMatrices.jl:36-60 exports); every device operation is one
call into libmodp_hip.so through the C ABI of include/modp.h.

    CuModMatrix(A, N; mod, new_size, elem_type)   CuModMatrix.jl:143-145
    zeros / eye / rand                            CuModMatrix.jl:510-556
    Array(A), size, rows, cols                    CuModMatrix.jl:220-261
    A * B, mul!(C, A, B)                          mul_ops.jl:54-58, CuModMatrix.jl:767-787
    stripe_mul!, find_max_stripe_ops              stripe_mul.jl:13-27, 175-244
    mat_mul_gpu_type, mat_mul_type_inplace!       mat_mul_gpu_direct.jl:8-85
    mod!, mod_elements!, change_modulus           mod_ops.jl:13-27, CuModMatrix.jl:712-760
    A ^ n                                         CuModMatrix.jl:307-329
    pluq_gpu_kernel(A) -> (U, L, P, Q)            rref_lu_pluq/pluq_kernels.jl:46-157
    mod_inv                                       pluq_kernels.jl:11-31
    apply_{row,col}[_inv]_perm!, perm_array_to_matrix   permutations.jl:11-157
    lu(A) -> (U, L, perm), rref(A)                test/Experiments/rref_gpu_type.jl:8-103

Julia's `!` suffix is spelled with a trailing underscore (mul_, apply_row_perm_).
"""
from __future__ import annotations

import ctypes
import weakref
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import FORM_AUTO, FORM_LIMB_I8, FORM_PLAIN_F64, FORM_RNS_I8, GemmJob, ElimJob
from ._lib import TUNE_CACHED, TUNE_OFF, TUNE_ON_MISS

TILE_WIDTH = 32
DEFAULT_TYPE = np.float32

__all__ = [
    "CuModArray", "CuModMatrix", "zeros", "eye", "rand", "Array", "rows", "cols",
    "mul_", "stripe_mul_", "find_max_stripe_ops", "mat_mul_gpu_type", "mat_mul_type_inplace_",
    "mod_", "mod_elements_", "change_modulus", "pluq_gpu_kernel", "pluq", "lu", "rref",
    "mod_inv", "apply_row_perm_", "apply_row_inv_perm_", "apply_col_perm_", "apply_col_inv_perm_",
    "perm_array_to_matrix", "gemm_batched", "GemmBatchPlan", "elim_batched",
    "CuModArraySizeMismatchException", "CuModArrayModulusMismatchException",
    "CuModMatrixNotSquareException", "CuModMatrixModulusNotPrimeException", "ArgumentError",
    "DimensionMismatch", "FORM_AUTO", "FORM_PLAIN_F64", "FORM_LIMB_I8", "FORM_RNS_I8", "context",
]


# ---------------------------------------------------------------------------
# exceptions (CuModMatrix.jl:5-31, plus Julia's ArgumentError/DimensionMismatch)
# ---------------------------------------------------------------------------

class CuModArraySizeMismatchException(Exception):
    pass


class CuModArrayModulusMismatchException(Exception):
    pass


class CuModMatrixNotSquareException(Exception):
    pass


class CuModMatrixModulusNotPrimeException(Exception):
    pass


class MatrixNotInvertibleException(Exception):
    """MatrixNotInvertibleException (raised by inverse, CuModMatrix.jl:485)."""


class InverseNotDefinedException(Exception):
    """InverseNotDefinedException (CuModMatrix.jl:29; raised by
    lower_triangular_inverse_no_copy, triangular_inverse_no_copy.jl:463)."""


class ArgumentError(ValueError):
    pass


class DimensionMismatch(ValueError):
    pass


class DeviceError(RuntimeError):
    pass


_STATUS_EXC = {
    _lib.MODP_ERR_DIM: CuModArraySizeMismatchException,
    _lib.MODP_ERR_MODULUS_RANGE: ArgumentError,
    _lib.MODP_ERR_MODULUS_MISMATCH: CuModArrayModulusMismatchException,
    _lib.MODP_ERR_NOT_INVERTIBLE: CuModMatrixModulusNotPrimeException,
    _lib.MODP_ERR_DTYPE: ArgumentError,
    _lib.MODP_ERR_OOM: MemoryError,
    _lib.MODP_ERR_DEVICE: DeviceError,
    _lib.MODP_ERR_ARG: ArgumentError,
    _lib.MODP_ERR_NO_DEVICE: DeviceError,
}


# ---------------------------------------------------------------------------
# context and device buffers
# ---------------------------------------------------------------------------

class Context:
    """One device, one HIP stream, the library's workspaces."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        st = self.lib.modp_ctx_create(device, ctypes.byref(h))
        if st != 0:
            raise DeviceError(f"modp_ctx_create(device={device}) failed with status {st} "
                              "(needs an MI355X / gfx950 GPU)")
        self.h = h
        self.device = device
        self._fin = weakref.finalize(self, self.lib.modp_ctx_destroy, h)

    def check(self, st: int, what: str = ""):
        if st == 0:
            return
        msg = self.lib.modp_last_error(self.h).decode(errors="replace")
        exc = _STATUS_EXC.get(st, DeviceError)
        raise exc(f"{what}: {msg}" if what else msg)

    def synchronize(self):
        self.check(self.lib.modp_synchronize(self.h), "synchronize")

    def malloc(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)

    # timing helpers used by the benchmark
    def kernel_timing(self, enable: bool):
        self.check(self.lib.modp_kernel_timing(self.h, 1 if enable else 0))

    def kernel_timing_reset(self):
        self.check(self.lib.modp_kernel_timing_reset(self.h))

    def kernel_time(self, prefix: str) -> Tuple[float, int]:
        tot = ctypes.c_double()
        cnt = ctypes.c_int64()
        self.check(self.lib.modp_kernel_time(self.h, prefix.encode(), ctypes.byref(tot), ctypes.byref(cnt)))
        return tot.value, cnt.value

    def event_record(self, slot: int):
        self.check(self.lib.modp_event_record(self.h, slot))

    def event_elapsed(self, a: int, b: int) -> float:
        ms = ctypes.c_float()
        self.check(self.lib.modp_event_elapsed(self.h, a, b, ctypes.byref(ms)))
        return ms.value

    def probe_peak(self, kind: int) -> float:
        """Measured peak: 0 int8 MFMA ops/s, 1 FP64 MFMA flop/s, 2 HBM bytes/s."""
        r = ctypes.c_double()
        self.check(self.lib.modp_probe_peak(self.h, int(kind), ctypes.byref(r)), "probe_peak")
        return r.value

    # autotuner (no counterpart in the reference)
    def autotune_mode(self, mode: int):
        """TUNE_OFF (heuristics), TUNE_CACHED, or TUNE_ON_MISS."""
        self.check(self.lib.modp_autotune_mode(self.h, int(mode)), "autotune_mode")

    def autotune_gemm(self, m: int, n: int, k: int, N: int, elem_type=np.float64):
        """Time every valid launch configuration on synthetic operands; the
        fastest is cached and returned as (config dict, milliseconds)."""
        cfg = _lib.GemmConfig()
        ms = ctypes.c_double()
        self.check(self.lib.modp_autotune_gemm(self.h, _dtype_code(elem_type), m, n, k, int(N), ctypes.byref(cfg),
                                               ctypes.byref(ms)), "autotune_gemm")
        return _cfg_dict(cfg), ms.value

    def gemm_config(self, m: int, n: int, k: int, N: int, elem_type=np.float64):
        """(config dict, tuned?) that a FORM_AUTO product of this shape runs."""
        cfg = _lib.GemmConfig()
        tuned = ctypes.c_int()
        self.check(self.lib.modp_gemm_config_get(self.h, _dtype_code(elem_type), m, n, k, int(N), ctypes.byref(cfg),
                                                 ctypes.byref(tuned)), "gemm_config")
        return _cfg_dict(cfg), bool(tuned.value)

    def autotune_export(self) -> str:
        n = ctypes.c_size_t()
        self.check(self.lib.modp_autotune_export(self.h, None, 0, ctypes.byref(n)), "autotune_export")
        buf = ctypes.create_string_buffer(n.value)
        self.check(self.lib.modp_autotune_export(self.h, buf, n.value, ctypes.byref(n)), "autotune_export")
        return buf.value.decode()

    def autotune_import(self, text: str):
        self.check(self.lib.modp_autotune_import(self.h, text.encode()), "autotune_import")


def _cfg_dict(cfg) -> dict:
    return {"form": cfg.form, "tile_m": cfg.tile_m, "tile_n": cfg.tile_n, "ksplit": cfg.ksplit}


_CTX: Optional[Context] = None


def context(device: int = 0) -> Context:
    """The process-wide default context (created on first use)."""
    global _CTX
    if _CTX is None:
        _CTX = Context(device)
    return _CTX


class DeviceBuffer:
    """Device allocation owned by the library's context, freed on GC."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        ctx.check(ctx.lib.modp_malloc(ctx.h, max(1, self.nbytes), ctypes.byref(p)), "malloc")
        self.ptr = p.value
        self._fin = weakref.finalize(self, ctx.lib.modp_free, ctx.h, ctypes.c_void_p(self.ptr))

    def upload(self, host: np.ndarray):
        host = np.ascontiguousarray(host) if not host.flags.f_contiguous else host
        assert host.nbytes <= self.nbytes
        self.ctx.check(self.ctx.lib.modp_memcpy_h2d(self.ctx.h, self.ptr, host.ctypes.data, host.nbytes), "upload")

    def download(self, host: np.ndarray):
        assert host.nbytes <= self.nbytes
        self.ctx.check(self.ctx.lib.modp_memcpy_d2h(self.ctx.h, host.ctypes.data, self.ptr, host.nbytes), "download")

    def zero(self):
        self.ctx.check(self.ctx.lib.modp_memset_zero(self.ctx.h, self.ptr, self.nbytes), "memset")


def _dtype_code(t) -> int:
    t = np.dtype(t)
    if t == np.float64:
        return _lib.MODP_F64
    if t == np.float32:
        return _lib.MODP_F32
    raise ArgumentError(f"Element type {t} unsupported by the device path (Float64/Float32)")


# ---------------------------------------------------------------------------
# CuModArray (CuModMatrix.jl:42-122)
# ---------------------------------------------------------------------------

class CuModArray:
    """A matrix mod N in padded column-major device storage.

    ``data`` is the device buffer of the padded array (rows+32) x (cols+32),
    ``size`` the logical size, ``N`` the modulus (CuModMatrix.jl:42-45)."""

    def __init__(self, data: DeviceBuffer, padded: Tuple[int, int], size: Tuple[int, int], N: int, elem_type):
        self.data = data
        self.padded = (int(padded[0]), int(padded[1]))
        self.size = (int(size[0]), int(size[1]))
        self.N = int(N)
        self.elem_type = np.dtype(elem_type)

    # -- storage -----------------------------------------------------------
    @property
    def ptr(self) -> int:
        return self.data.ptr

    @property
    def ld(self) -> int:
        return self.padded[0]

    @property
    def dtype_code(self) -> int:
        return _dtype_code(self.elem_type)

    @property
    def ctx(self) -> Context:
        return self.data.ctx

    def shape(self):
        return self.size

    def __len__(self):
        return self.size[0] * self.size[1]

    def eltype(self):
        return self.elem_type

    def padded_array(self) -> np.ndarray:
        """unsafe_Array: the whole padded storage (CuModMatrix.jl:251-253)."""
        out = np.empty(self.padded, dtype=self.elem_type, order="F")
        self.data.download(out)
        return out

    def __array__(self, dtype=None):
        a = Array(self)
        return a if dtype is None else a.astype(dtype)

    def __repr__(self):
        return f"{self.size[0]}×{self.size[1]} CuModMatrix modulo {self.N}:\n{Array(self)}"

    # -- algebra -------------------------------------------------------------
    def __mul__(self, other):
        if isinstance(other, CuModArray):
            return _matmul(self, other)
        return NotImplemented

    __matmul__ = __mul__

    def __pow__(self, n: int):
        return _power(self, int(n))


CuModMatrix_t = CuModArray


def _pad(d: int) -> int:
    return d + TILE_WIDTH


def _new(rows_: int, cols_: int, N: int, elem_type, ctx: Optional[Context] = None, zero: bool = True,
         padded: Optional[Tuple[int, int]] = None) -> CuModArray:
    ctx = ctx or context()
    pr, pc = padded or (_pad(rows_), _pad(cols_))
    buf = ctx.malloc(pr * pc * np.dtype(elem_type).itemsize)
    if zero:
        buf.zero()
    return CuModArray(buf, (pr, pc), (rows_, cols_), N, elem_type)


def CuModMatrix(A, N: int, mod: bool = True, new_size=None, elem_type=DEFAULT_TYPE,
                ctx: Optional[Context] = None) -> CuModArray:
    """Construct from a host matrix (CuModMatrix.jl:53-99): pad by +32 per
    dimension with zeros, convert (inexact conversion raises), copy to the
    device and reduce mod N."""
    N = int(N)
    if N > 2 ** 52:
        raise CuModArrayModulusMismatchException(
            f"Modulus {N} is bigger than 2^52 = {2 ** 52}, the largest supported modulus")
    host = np.asarray(A)
    if host.ndim == 1:
        host = host.reshape(-1, 1)
    if host.ndim != 2:
        raise ArgumentError("CuModMatrix expects a matrix")
    et = np.dtype(elem_type)
    _dtype_code(et)
    conv = host.astype(et)
    if host.size and not np.array_equal(conv.astype(np.float64), host.astype(np.float64)):
        raise ArgumentError(f"InexactError: cannot convert input to {et} exactly")
    r, c = host.shape
    M = _new(r, c, N, et, ctx, zero=False)
    padded = np.zeros(M.padded, dtype=et, order="F")
    padded[:r, :c] = conv
    M.data.upload(padded)
    if mod and r and c:
        mod_(M, M, N)
    if new_size is not None:
        M.size = (int(new_size[0]), int(new_size[1]))
    return M


CuModArray_from_host = CuModMatrix


def rows(A: CuModArray) -> int:
    return A.size[0]


def cols(A: CuModArray) -> int:
    return A.size[1]


def size(A: CuModArray, i: Optional[int] = None):
    return A.size if i is None else A.size[i - 1]


def Array(A: CuModArray) -> np.ndarray:
    """Copy to host without the padding (CuModMatrix.jl:256-261)."""
    full = A.padded_array()
    return np.array(full[:A.size[0], :A.size[1]], order="F")


def CuModVector(x, N: int, mod: bool = True, elem_type=DEFAULT_TYPE, ctx: Optional[Context] = None) -> CuModArray:
    """A length-n vector mod N, held as an n x 1 CuModArray (the column of a
    CuModVector, CuModMatrix.jl:42-122); `mul_(z, A, x)` with such vectors is
    the matrix-vector product of stripe_mul!(z, A, x) (stripe_mul.jl:82-168)."""
    v = np.asarray(x).reshape(-1, 1)
    return CuModMatrix(v, N, mod=mod, elem_type=elem_type, ctx=ctx)


def zeros(T, rows_: int, cols_: int, N: int, new_size=None, ctx: Optional[Context] = None) -> CuModArray:
    """zeros(T, rows, cols, N) (CuModMatrix.jl:530-534)."""
    M = _new(rows_, cols_, N, T, ctx)
    if new_size is not None:
        M.size = tuple(new_size)
    return M


def eye(T, n: int, N: int, ctx: Optional[Context] = None) -> CuModArray:
    """eye(T, n, N) (CuModMatrix.jl:510-513)."""
    M = _new(n, n, N, T, ctx)
    M.ctx.check(M.ctx.lib.modp_eye(M.ctx.h, M.dtype_code, n, n, M.ptr, M.ld), "eye")
    return M


def rand(T, rows_: int, cols_: int, N: int, seed: int = 0, offset: int = 0,
         ctx: Optional[Context] = None) -> CuModArray:
    """Seeded uniform residues (stands in for rand(T,rows,cols,N),
    CuModMatrix.jl:551-556, which is unseeded and leaves nonzero padding; here
    the padding stays zero).  Element (i,j) is draw offset + i + j*rows of the
    counter-based SplitMix64 stream `seed`, Lemire-mapped to [0,N)."""
    M = _new(rows_, cols_, N, T, ctx)
    M.ctx.check(M.ctx.lib.modp_rand(M.ctx.h, M.dtype_code, rows_, cols_, M.ptr, M.ld, int(N),
                                    ctypes.c_uint64(seed), ctypes.c_uint64(offset)), "rand")
    return M


def mod_(C: CuModArray, A: CuModArray, m: int) -> CuModArray:
    """mod!(C, A, m) (mod_ops.jl:13-27) on the padded storage."""
    if C is not A:
        if C.padded != A.padded or C.elem_type != A.elem_type:
            raise CuModArraySizeMismatchException("mod!: storage shapes differ")
        C.ctx.check(C.ctx.lib.modp_memcpy_d2d(C.ctx.h, C.ptr, A.ptr, A.data.nbytes), "copy")
    C.ctx.check(C.ctx.lib.modp_mod(C.ctx.h, C.dtype_code, C.padded[0], C.padded[1], C.ptr, C.ld, int(m)), "mod!")
    return C


def mod_elements_(A: CuModArray, mod_N: int = -1) -> CuModArray:
    """mod_elements!(A, [mod_N]) (CuModMatrix.jl:712-716)."""
    return mod_(A, A, mod_N if mod_N > 0 else A.N)


def change_modulus(A: CuModArray, new_N: int) -> CuModArray:
    """change_modulus(A, new_N) (CuModMatrix.jl:730-741)."""
    R = _new(A.size[0], A.size[1], new_N, A.elem_type, A.ctx, zero=False, padded=A.padded)
    R.ctx.check(R.ctx.lib.modp_memcpy_d2d(R.ctx.h, R.ptr, A.ptr, A.data.nbytes), "copy")
    if new_N < A.N:
        mod_(R, R, new_N)
    return R


def change_modulus_no_alloc_(A: CuModArray, new_N: int) -> CuModArray:
    """change_modulus_no_alloc!(A, new_N) (CuModMatrix.jl:751-758): A's
    storage reduced in place when new_N < A.N, returned under modulus new_N."""
    if new_N < A.N:
        mod_(A, A, new_N)
    return CuModArray(A.data, A.padded, A.size, new_N, A.elem_type)


def copy(A: CuModArray) -> CuModArray:
    """CuModcopy (CuModMatrix.jl:838-840)."""
    R = _new(A.size[0], A.size[1], A.N, A.elem_type, A.ctx, zero=False, padded=A.padded)
    R.ctx.check(R.ctx.lib.modp_memcpy_d2d(R.ctx.h, R.ptr, A.ptr, A.data.nbytes), "copy")
    return R


# ---------------------------------------------------------------------------
# GEMM
# ---------------------------------------------------------------------------

def find_max_stripe_ops(elem_type, N: int) -> int:
    """stripe_mul.jl:13-27: floor((2^bits - 1)/(N-1)^2) - 1, bits 53/24/11."""
    t = np.dtype(elem_type)
    bits = {np.dtype(np.float64): 53, np.dtype(np.float32): 24, np.dtype(np.float16): 11}.get(t)
    if bits is None:
        raise ArgumentError("The input type is not supported for CUBLAS gemm")
    if N == 1:
        return 2 ** 63 - 1
    return (2 ** bits - 1) // ((N - 1) ** 2) - 1


def _gemm_raw(C: CuModArray, A: CuModArray, B: CuModArray, N: int, form: int = FORM_AUTO, config=None):
    m, k = A.size
    n = B.size[1]
    ctx = C.ctx
    if n == 1 and form == FORM_AUTO and config is None:
        # matrix-vector product: one streaming pass over A
        ctx.check(ctx.lib.modp_gemv(ctx.h, C.dtype_code, m, k, A.ptr, A.ld, B.ptr, C.ptr, int(N)), "gemv")
        return
    if config is not None:
        cfg = _lib.GemmConfig(**dict(config))
        ctx.check(ctx.lib.modp_gemm_ex(ctx.h, ctypes.byref(cfg), C.dtype_code, m, n, k, A.ptr, A.ld, B.ptr, B.ld,
                                       C.ptr, C.ld, int(N)), "gemm")
        return
    ctx.check(ctx.lib.modp_gemm(ctx.h, form, C.dtype_code, m, n, k, A.ptr, A.ld, B.ptr, B.ld, C.ptr, C.ld, int(N)),
              "gemm")


def stripe_mul_(C: CuModArray, A: CuModArray, B: CuModArray, M=None, N=None, form: int = FORM_AUTO,
                config=None) -> None:
    """stripe_mul!(C, A, B; M, N) (stripe_mul.jl:175-244): C = A*B mod N.

    The device kernels keep every partial sum exact for any N <= 2^26, so
    the stripe width M is only used for the reference's range check.
    ``config`` (a dict with keys form, tile_m, tile_n, ksplit) forces a
    launch configuration instead of the form / autotuner choice."""
    if A.N != B.N or B.N != C.N:
        raise ArgumentError("Mismatched modulus in matmul")
    if cols(A) != rows(B):
        raise DimensionMismatch("")
    if A.elem_type != B.elem_type or B.elem_type != C.elem_type:
        raise ArgumentError("Mismatched element types in matmul")
    if A.elem_type not in (np.dtype(np.float64), np.dtype(np.float32)):
        raise ArgumentError(f"Element type {A.elem_type} unsupported by CUBLAS")
    R = A.N if N is None else int(N)
    if M is None:
        M = find_max_stripe_ops(A.elem_type, R)
    if M < 1:
        raise ArgumentError(f"cannot perform a single multiplication for modulus {A.N} with datatype "
                            f"{'Float64' if A.elem_type == np.float64 else 'Float32'}")
    _gemm_raw(C, A, B, C.N if N is None else R, form, config)


def mul_(C: CuModArray, A: CuModArray, B: CuModArray, M=None, N=None, form: int = FORM_AUTO,
         config=None) -> CuModArray:
    """LinearAlgebra.mul!(C, A, B; M, N) (CuModMatrix.jl:767-787)."""
    if A.N != B.N or A.N != C.N:
        raise CuModArrayModulusMismatchException("All matrices must have the same modulus N")
    if cols(A) != rows(B):
        raise CuModArraySizeMismatchException("Matrix dimensions do not match for multiplication")
    if rows(C) != rows(A) or cols(C) != cols(B):
        raise CuModArraySizeMismatchException("Output matrix C has incorrect dimensions")
    stripe_mul_(C, A, B, M=M, N=N, form=form, config=config)
    return C


def matmul(A: CuModArray, B: CuModArray, form: int = FORM_AUTO) -> CuModArray:
    """Base.:*(A, B) (mul_ops.jl:54-58): allocate C with zeros, stripe_mul!."""
    C = zeros(A.elem_type, rows(A), cols(B), A.N, ctx=A.ctx)
    stripe_mul_(C, A, B, form=form)
    return C


_matmul = matmul


def mat_mul_gpu_type(A: CuModArray, B: CuModArray, mod_N: int = -1, form: int = FORM_AUTO) -> CuModArray:
    """mat_mul_gpu_type(A, B, [mod_N]) (mat_mul_gpu_direct.jl:8-42)."""
    N = mod_N if mod_N > 0 else A.N
    if cols(A) != rows(B):
        raise CuModArraySizeMismatchException(
            f"Matrix dimensions do not match. A has {rows(A)} rows and {cols(A)} cols, "
            f"B has {rows(B)} rows and {cols(B)} cols.")
    C = zeros(A.elem_type, rows(A), cols(B), N, ctx=A.ctx)
    _gemm_raw(C, A, B, N, form)
    return C


def mat_mul_type_inplace_(C: CuModArray, A: CuModArray, B: CuModArray, mod_N: int = -1,
                          form: int = FORM_AUTO) -> CuModArray:
    """mat_mul_type_inplace!(C, A, B, [mod_N]) (mat_mul_gpu_direct.jl:49-85)."""
    N = mod_N if mod_N > 0 else A.N
    if rows(C) != rows(A) or cols(C) != cols(B):
        raise CuModArraySizeMismatchException("Output matrix C has incorrect dimensions")
    if cols(A) != rows(B):
        raise CuModArraySizeMismatchException("Matrix dimensions do not match")
    _gemm_raw(C, A, B, N, form)
    return C


def _power(A: CuModArray, n: int) -> CuModArray:
    """Base.:^(A, n) by recursive squaring (CuModMatrix.jl:307-329)."""
    if rows(A) != cols(A):
        raise CuModMatrixNotSquareException("Matrix must be square for power operation")
    if n == 0:
        return CuModMatrix(np.eye(rows(A)), A.N, elem_type=A.elem_type, ctx=A.ctx)
    if n < 0:
        raise CuModMatrixModulusNotPrimeException("negative powers need the inverse, which is out of scope")
    if n == 1:
        return A
    if n % 2 == 0:
        h = _power(A, n // 2)
        return matmul(h, h)
    return matmul(A, _power(A, n - 1))


# ---------------------------------------------------------------------------
# elimination
# ---------------------------------------------------------------------------

def mod_inv(p: int, P: int) -> int:
    """pluq_kernels.jl:11-31 (extended Euclid, result in [0, P))."""
    inv, new_inv, rem, new_rem = 0, 1, P, p
    while new_rem != 0:
        q = rem // new_rem
        inv, new_inv = new_inv, inv - q * new_inv
        rem, new_rem = new_rem, rem - q * new_rem
    if inv < 0:
        inv += P
    return inv


def _stack(buf: np.ndarray, n: int) -> List[Tuple[int, int]]:
    return [(int(buf[2 * i]), int(buf[2 * i + 1])) for i in range(n)]


def pluq(A: CuModArray, form: int = FORM_AUTO):
    """P*A*Q = L*U mod N.  Returns (U, L, Perm_rows, Perm_cols, rank)."""
    m, n = A.size
    U = zeros(A.elem_type, m, n, A.N, ctx=A.ctx)
    L = zeros(A.elem_type, m, m, A.N, ctx=A.ctx)
    rs = np.zeros(2 * max(1, min(m, n)), dtype=np.int64)
    cs = np.zeros(4 * max(1, n), dtype=np.int64)
    nr, nc, rk = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    p64 = ctypes.POINTER(ctypes.c_int64)
    ctx = A.ctx
    ctx.check(ctx.lib.modp_pluq(ctx.h, form, A.dtype_code, m, n, A.ptr, A.ld, U.ptr, U.ld, L.ptr, L.ld, A.N,
                                rs.ctypes.data_as(p64), ctypes.byref(nr), cs.ctypes.data_as(p64),
                                ctypes.byref(nc), ctypes.byref(rk)), "pluq")
    return U, L, _stack(rs, nr.value), _stack(cs, nc.value), rk.value


def crt(A1: CuModArray, A2: CuModArray) -> CuModArray:
    """Chinese remaindering of two residue matrices of equal size: the
    CuModArray of modulus N1*N2 whose entries reduce to A1 mod N1 and A2 mod N2
    (the composite modulus of KaratsubaMatrix.jl:18-51, built from per-prime
    results)."""
    if A1.size != A2.size:
        raise CuModArraySizeMismatchException("CRT needs matrices of the same size")
    if A1.elem_type != A2.elem_type:
        raise ArgumentError("Mismatched element types")
    m, n = A1.size
    N = A1.N * A2.N
    Y = zeros(A1.elem_type, m, n, N, ctx=A1.ctx)
    ctx = A1.ctx
    ctx.check(ctx.lib.modp_crt2(ctx.h, A1.dtype_code, m, n, A1.ptr, A1.ld, A1.N, A2.ptr, A2.ld, A2.N, Y.ptr, Y.ld),
              "crt")
    return Y


def crt_words(moduli: Sequence[int], signed: bool = False) -> int:
    """32-bit words needed for the product of the moduli (and a sign bit)."""
    import math
    bits = sum(math.log2(int(q)) for q in moduli) + (1 if signed else 0)
    return int(math.ceil(bits / 32 - 1e-12))


def crt_reconstruct(mats: Sequence[CuModArray], signed: bool = False) -> np.ndarray:
    """Chinese remaindering of residue matrices of equal size under pairwise
    coprime moduli (their N) into the integer matrix y with y = A_i mod N_i:
    y in [0, P) or, signed, in (-P/2, P/2], P = prod N_i.  Returns a numpy
    object array of Python integers (the multi-word device result
    reassembled on the host)."""
    if not mats:
        raise ArgumentError("CRT needs at least one residue matrix")
    m, n = mats[0].size
    et = mats[0].elem_type
    for A in mats:
        if A.size != (m, n):
            raise CuModArraySizeMismatchException("CRT needs matrices of the same size")
        if A.elem_type != et:
            raise ArgumentError("Mismatched element types")
    ctx = mats[0].ctx
    r = len(mats)
    moduli = [int(A.N) for A in mats]
    words = crt_words(moduli, signed)
    out = np.zeros(max(1, words * m * n), dtype=np.uint32)
    buf = ctx.malloc(4 * out.size)
    ptrs = (ctypes.c_void_p * r)(*[A.ptr for A in mats])
    lds = (ctypes.c_int64 * r)(*[A.ld for A in mats])
    ps = (ctypes.c_int64 * r)(*moduli)
    ctx.check(ctx.lib.modp_crt(ctx.h, mats[0].dtype_code, m, n, r, ptrs, lds, ps, 1 if signed else 0, buf.ptr,
                               max(m, 1), words), "crt")
    buf.download(out)
    planes = out[:words * m * n].reshape(words, n, m)  # plane w: column-major m x n
    Y = np.zeros((m, n), dtype=object)
    for w in range(words - 1, -1, -1):
        Y = Y * (1 << 32) + planes[w].T.astype(object)
    if signed:
        top = 1 << (32 * words)
        Y = np.where(Y >= top // 2, Y - top, Y)
    return Y


def _trtri(A: CuModArray, uplo: int) -> CuModArray:
    m, n = A.size
    Ainv = zeros(A.elem_type, n, m, A.N, ctx=A.ctx)
    ctx = A.ctx
    st = ctx.lib.modp_trtri(ctx.h, A.dtype_code, uplo, m, n, A.ptr, A.ld, Ainv.ptr, Ainv.ld, A.N)
    if st == _lib.MODP_ERR_DIM:
        raise InverseNotDefinedException(ctx.lib.modp_last_error(ctx.h).decode(errors="replace"))
    if st == _lib.MODP_ERR_NOT_INVERTIBLE:
        raise MatrixNotInvertibleException(ctx.lib.modp_last_error(ctx.h).decode(errors="replace"))
    ctx.check(st, "triangular inverse")
    return Ainv


def upper_triangular_inverse_no_copy(A: CuModArray, debug: bool = False) -> CuModArray:
    """upper_triangular_inverse_no_copy(A) (triangular_inverse_no_copy.jl:197-229):
    the (right) inverse of an upper triangular matrix, size (cols, rows)."""
    return _trtri(A, 0)


def lower_triangular_inverse_no_copy(A: CuModArray, debug: bool = False) -> CuModArray:
    """lower_triangular_inverse_no_copy(A) (triangular_inverse_no_copy.jl:450-478);
    a tall lower-triangular matrix raises InverseNotDefinedException."""
    return _trtri(A, 1)


def _inverse(A: CuModArray, form: int):
    m, n = A.size
    if m != n:
        raise CuModMatrixNotSquareException("Matrix must be square to be inverted")
    Ainv = zeros(A.elem_type, n, n, A.N, ctx=A.ctx)
    ok = ctypes.c_int()
    ctx = A.ctx
    ctx.check(ctx.lib.modp_inverse(ctx.h, form, A.dtype_code, n, A.ptr, A.ld, Ainv.ptr, Ainv.ld, A.N,
                                   ctypes.byref(ok)), "inverse")
    return bool(ok.value), Ainv


def inverse(A: CuModArray, form: int = FORM_AUTO) -> CuModArray:
    """inverse(A) (CuModMatrix.jl:480-502): A^{-1} mod N from PLUQ and the two
    triangular inverses; raises MatrixNotInvertibleException when the mod-N
    rank is deficient."""
    ok, Ainv = _inverse(A, form)
    if not ok:
        raise MatrixNotInvertibleException(f"Matrix is not invertible mod {A.N}")
    return Ainv


def is_invertible_with_inverse(A: CuModArray, debug: bool = False, form: int = FORM_AUTO):
    """is_invertible_with_inverse(A) (CuModMatrix.jl:351-411): (true, A^{-1})
    or (false, None)."""
    ok, Ainv = _inverse(A, form)
    return (True, Ainv) if ok else (False, None)


def is_invertible(A: CuModArray, form: int = FORM_AUTO) -> bool:
    """is_invertible(A) (CuModMatrix.jl:460-465): full rank mod N, from PLUQ."""
    return pluq(A, form=form)[4] == min(A.size)


def pluq_gpu_kernel(A: CuModArray, debug: bool = False, form: int = FORM_AUTO):
    """pluq_gpu_kernel(A; debug) -> (U, L, Perm_rows, Perm_cols)
    (pluq_kernels.jl:46-157).  A is not modified."""
    U, L, P, Q, _ = pluq(A, form=form)
    return U, L, P, Q


def lu(A: CuModArray, form: int = FORM_AUTO):
    """PLE: returns (U, L, perm) with A[perm, :] = L*U (1-based perm), as the
    prototype lu_gpu_type (test/Experiments/rref_gpu_type.jl:60-103)."""
    U, L, perm, _, _ = lu_full(A, form)
    return U, L, perm


def lu_full(A: CuModArray, form: int = FORM_AUTO):
    m, n = A.size
    U = zeros(A.elem_type, m, n, A.N, ctx=A.ctx)
    L = zeros(A.elem_type, m, m, A.N, ctx=A.ctx)
    perm = np.zeros(max(1, m), dtype=np.int64)
    piv = np.zeros(max(1, min(m, n)), dtype=np.int64)
    rk = ctypes.c_int64()
    p64 = ctypes.POINTER(ctypes.c_int64)
    ctx = A.ctx
    ctx.check(ctx.lib.modp_lu(ctx.h, form, A.dtype_code, m, n, A.ptr, A.ld, U.ptr, U.ld, L.ptr, L.ld, A.N,
                              perm.ctypes.data_as(p64), piv.ctypes.data_as(p64), ctypes.byref(rk)), "lu")
    r = rk.value
    return U, L, [int(x) for x in perm[:m]], [int(x) for x in piv[:r]], r


def rref(A: CuModArray, form: int = FORM_AUTO) -> CuModArray:
    """Reduced row echelon form mod N (rref_gpu_type, rref_gpu_type.jl:8-51)."""
    R, _, _ = rref_full(A, form)
    return R


def rref_full(A: CuModArray, form: int = FORM_AUTO):
    m, n = A.size
    R = zeros(A.elem_type, m, n, A.N, ctx=A.ctx)
    piv = np.zeros(max(1, min(m, n)), dtype=np.int64)
    rk = ctypes.c_int64()
    p64 = ctypes.POINTER(ctypes.c_int64)
    ctx = A.ctx
    ctx.check(ctx.lib.modp_rref(ctx.h, form, A.dtype_code, m, n, A.ptr, A.ld, R.ptr, R.ld, A.N,
                                piv.ctypes.data_as(p64), ctypes.byref(rk)), "rref")
    r = rk.value
    return R, [int(x) for x in piv[:r]], r


# ---------------------------------------------------------------------------
# permutations (permutations.jl)
# ---------------------------------------------------------------------------

def _perm(fn, P: Sequence[Tuple[int, int]], A: CuModArray, inverse: int, rowwise: bool):
    st = np.asarray(P, dtype=np.int64).reshape(-1)
    ctx = A.ctx
    # the reference kernels sweep the padded extent (permutations.jl:15,77)
    r, c = A.padded
    f = getattr(ctx.lib, fn)
    ctx.check(f(ctx.h, A.dtype_code, r, c, A.ptr, A.ld, st.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                len(P), inverse), fn)


def apply_row_perm_(P, A: CuModArray):
    """apply_row_perm!(P, A) (permutations.jl:73-80)."""
    _perm("modp_apply_row_perm", P, A, 0, True)


def apply_row_inv_perm_(P, A: CuModArray):
    """apply_row_inv_perm!(P, A) (permutations.jl:82-89)."""
    _perm("modp_apply_row_perm", P, A, 1, True)


def apply_col_perm_(P, A: CuModArray):
    """apply_col_perm!(P, A) (permutations.jl:11-18)."""
    _perm("modp_apply_col_perm", P, A, 0, False)


def apply_col_inv_perm_(P, A: CuModArray):
    """apply_col_inv_perm!(P, A) (permutations.jl:20-27)."""
    _perm("modp_apply_col_perm", P, A, 1, False)


def perm_array_to_matrix(perm, N: int, new_size, perm_stack: bool = False, elem_type=DEFAULT_TYPE) -> CuModArray:
    """perm_array_to_matrix (permutations.jl:141-157)."""
    n = len(perm)
    if perm_stack:
        P = np.eye(n, dtype=np.int64)
        for i, j in perm:
            P[[i - 1, j - 1], :] = P[[j - 1, i - 1], :]
    else:
        P = np.zeros((n, n), dtype=np.int64)
        for i in range(n):
            P[perm[i] - 1, i] = 1
    return CuModMatrix(P, N, new_size=new_size, elem_type=elem_type)


# ---------------------------------------------------------------------------
# batched multi-modular forms
# ---------------------------------------------------------------------------

class GemmBatchPlan:
    """A captured batch of independent products C_i = A_i*B_i mod p_i."""

    def __init__(self, jobs: Sequence[Tuple[CuModArray, CuModArray, CuModArray, int]], elem_type=np.float64):
        self.ctx = jobs[0][0].ctx if jobs else context()
        arr = (GemmJob * max(1, len(jobs)))()
        for i, (A, B, C, p) in enumerate(jobs):
            if cols(A) != rows(B) or rows(C) != rows(A) or cols(C) != cols(B):
                raise CuModArraySizeMismatchException(f"job {i}: dimensions do not match")
            arr[i] = GemmJob(rows(A), cols(B), cols(A), int(p), A.ptr, A.ld, B.ptr, B.ld, C.ptr, C.ld)
        self._keep = jobs
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.modp_gemm_batched_plan(self.ctx.h, _dtype_code(elem_type), arr, len(jobs),
                                                           ctypes.byref(h)), "gemm_batched_plan")
        self.h = h
        self._fin = weakref.finalize(self, self.ctx.lib.modp_gemm_batched_destroy, self.ctx.h, h)

    def run(self):
        self.ctx.check(self.ctx.lib.modp_gemm_batched_run(self.ctx.h, self.h), "gemm_batched_run")


def gemm_batched(jobs, elem_type=np.float64):
    """Run independent products C_i = A_i*B_i mod p_i in one persistent launch."""
    plan = GemmBatchPlan(jobs, elem_type)
    plan.run()
    plan.ctx.synchronize()


class ElimBatchPlan:
    """A captured batch of independent eliminations (ElimJob descriptors over
    device memory); run() is asynchronous on the context's stream."""

    def __init__(self, op: int, jobs_raw: Sequence[ElimJob], elem_type=np.float64, ctx: Optional[Context] = None):
        self.ctx = ctx or context()
        arr = (ElimJob * max(1, len(jobs_raw)))(*jobs_raw)
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.modp_elim_batched_plan(self.ctx.h, int(op), _dtype_code(elem_type), arr,
                                                           len(jobs_raw), ctypes.byref(h)), "elim_batched_plan")
        self.h = h
        self._fin = weakref.finalize(self, self.ctx.lib.modp_elim_batched_destroy, self.ctx.h, h)

    def run(self):
        self.ctx.check(self.ctx.lib.modp_elim_batched_run(self.ctx.h, self.h), "elim_batched_run")


def elim_batched(op: int, jobs_raw: Sequence[ElimJob], elem_type=np.float64, ctx: Optional[Context] = None):
    """Batched eliminations over raw ElimJob descriptors (device pointers)."""
    ctx = ctx or context()
    arr = (ElimJob * max(1, len(jobs_raw)))(*jobs_raw)
    ctx.check(ctx.lib.modp_elim_batched(ctx.h, int(op), _dtype_code(elem_type), arr, len(jobs_raw)), "elim_batched")
