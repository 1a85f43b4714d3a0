# GPUFiniteFieldMatricesAMD.jl -- Julia binding of libmodp_hip.so (MI355X).
#
# Drop-in for the CuModMatrix hot path of GPUFiniteFieldMatrices.jl: the same
# type shape (padded column-major `data`, logical `size`, modulus `N`), the
# same method names (`*`, `mul!`, `stripe_mul!`, `pluq_gpu_kernel`, `Array`,
# `apply_row_perm!`, ...), and the same exceptions.  Every device operation is
# one `ccall` into the C ABI declared in include/modp.h.  No CUDA.jl.
#
# Julia is not installed where this library is built and tested; this file is
# the binding a maintainer loads next to the built library.
module GPUFiniteFieldMatricesAMD

using LinearAlgebra

export ModMatrix, CuModMatrix, pluq_gpu_kernel, pluq, lu_ple, lu_gpu_type, rref, rref_gpu_type, mod_inv,
       apply_row_perm!, apply_row_inv_perm!, apply_col_perm!, apply_col_inv_perm!,
       stripe_mul!, find_max_stripe_ops, mat_mul_gpu_type, mat_mul_type_inplace!,
       inverse, is_invertible, is_invertible_with_inverse,
       upper_triangular_inverse_no_copy, lower_triangular_inverse_no_copy, crt,
       eye, mod_elements!, change_modulus, change_modulus_no_alloc!, crt_reconstruct

const LIB = get(ENV, "MODP_HIP_LIB", joinpath(@__DIR__, "..", "gpufinitefieldmatrices", "libmodp_hip.so"))
const TILE_WIDTH = 32
const DEFAULT_TYPE = Float32

# ---- exceptions (src/CuModMatrix/CuModMatrix.jl:5-31) ---------------------
struct CuModArraySizeMismatchException <: Exception
    message::String
end
struct CuModArrayModulusMismatchException <: Exception
    message::String
end
struct CuModMatrixModulusNotPrimeException <: Exception
    message::String
end
struct CuModMatrixNotSquareException <: Exception
    message::String
end
struct InverseNotDefinedException <: Exception
    message::String
end
struct MatrixNotInvertibleException <: Exception
    message::String
end

const MODP_OK = 0
const FORM_AUTO, FORM_PLAIN_F64, FORM_LIMB_I8, FORM_RNS_I8 = 0, 1, 2, 3

# ---- context ----------------------------------------------------------------
mutable struct Context
    h::Ptr{Cvoid}
end

const CTX = Ref{Union{Nothing,Context}}(nothing)

function context()
    if CTX[] === nothing
        h = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:modp_ctx_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, h)
        st == MODP_OK || error("modp_ctx_create failed with status $st (needs an MI355X / gfx950)")
        c = Context(h[])
        finalizer(c -> ccall((:modp_ctx_destroy, LIB), Cint, (Ptr{Cvoid},), c.h), c)
        CTX[] = c
    end
    return CTX[]::Context
end

function check(st::Integer, what::AbstractString="")
    st == MODP_OK && return nothing
    msg = unsafe_string(ccall((:modp_last_error, LIB), Cstring, (Ptr{Cvoid},), context().h))
    m = isempty(what) ? msg : "$what: $msg"
    st == 1 && throw(CuModArraySizeMismatchException(m))
    st == 2 && throw(ArgumentError(m))
    st == 3 && throw(CuModArrayModulusMismatchException(m))
    st == 4 && throw(CuModMatrixModulusNotPrimeException(m))
    st == 5 && throw(ArgumentError(m))
    st == 6 && throw(OutOfMemoryError())
    error(m)
end

dtype_code(::Type{Float64}) = Cint(0)
dtype_code(::Type{Float32}) = Cint(1)
dtype_code(T) = throw(ArgumentError("Element type $T unsupported by the device path"))

# ---- device buffers -----------------------------------------------------------
mutable struct DeviceBuffer
    ptr::Ptr{Cvoid}
    nbytes::Int
end

function DeviceBuffer(nbytes::Integer)
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:modp_malloc, LIB), Cint, (Ptr{Cvoid}, Csize_t, Ref{Ptr{Cvoid}}), context().h, max(nbytes, 1), p), "malloc")
    b = DeviceBuffer(p[], nbytes)
    finalizer(b -> ccall((:modp_free, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), context().h, b.ptr), b)
    return b
end

# ---- the matrix type (CuModArray{T,2}, CuModMatrix.jl:42-122) -----------------
struct ModMatrix{T} <: AbstractMatrix{T}
    data::DeviceBuffer      # padded (rows+32) x (cols+32), column-major
    padded::Dims{2}
    size::Dims{2}
    N::Int
end

const CuModMatrix = ModMatrix

Base.size(A::ModMatrix) = A.size
Base.eltype(::ModMatrix{T}) where {T} = T
rows(A::ModMatrix) = A.size[1]
cols(A::ModMatrix) = A.size[2]
ld(A::ModMatrix) = A.padded[1]

function ModMatrix{T}(rows::Integer, cols::Integer, N::Integer; zero::Bool=true) where {T}
    pr, pc = rows + TILE_WIDTH, cols + TILE_WIDTH
    buf = DeviceBuffer(pr * pc * sizeof(T))
    zero && check(ccall((:modp_memset_zero, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), context().h, buf.ptr, buf.nbytes))
    return ModMatrix{T}(buf, (pr, pc), (rows, cols), N)
end

"""
    ModMatrix(A, N; mod=true, new_size=nothing, elem_type=Float32)

Constructor from a host matrix (CuModMatrix.jl:143-145, 53-99).
"""
function ModMatrix(A::AbstractMatrix, N::Integer; mod::Bool=true, new_size=nothing, elem_type=DEFAULT_TYPE)
    2^52 < N && throw(CuModArrayModulusMismatchException(
        "Modulus $N is bigger than 2^52 = $(2^52), the largest supported modulus"))
    T = elem_type
    M = ModMatrix{T}(size(A, 1), size(A, 2), N; zero=false)
    host = Base.zeros(T, M.padded...)
    host[1:size(A, 1), 1:size(A, 2)] .= convert.(T, A)   # errors when inexact
    check(ccall((:modp_memcpy_h2d, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                context().h, M.data.ptr, host, sizeof(host)), "upload")
    if mod
        check(ccall((:modp_mod, LIB), Cint, (Ptr{Cvoid}, Cint, Int64, Int64, Ptr{Cvoid}, Int64, Int64),
                    context().h, dtype_code(T), M.padded[1], M.padded[2], M.data.ptr, ld(M), N), "mod!")
    end
    new_size === nothing && return M
    return ModMatrix{T}(M.data, M.padded, Tuple(new_size), N)
end

function Base.Array(A::ModMatrix{T}) where {T}
    host = Matrix{T}(undef, A.padded...)
    check(ccall((:modp_memcpy_d2h, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                context().h, host, A.data.ptr, sizeof(host)), "download")
    return host[1:A.size[1], 1:A.size[2]]
end

# the package's own zeros(T, rows, cols, N) (CuModMatrix.jl:530-534): a new
# function of this module, so host arrays above use Base.zeros explicitly
zeros(::Type{T}, r::Integer, c::Integer, N::Integer) where {T} = ModMatrix{T}(r, c, N)

"""
    eye(T, n, N)

n x n identity mod N (CuModMatrix.jl:510-513).
"""
function eye(::Type{T}, n::Integer, N::Integer) where {T}
    M = ModMatrix{T}(n, n, N)
    check(ccall((:modp_eye, LIB), Cint, (Ptr{Cvoid}, Cint, Int64, Int64, Ptr{Cvoid}, Int64),
                context().h, dtype_code(T), n, n, M.data.ptr, ld(M)), "eye")
    return M
end

"""
    mod_elements!(A, mod_N=-1)

Reduce every stored entry (padding included) mod mod_N, or mod A.N when
mod_N <= 0 (CuModMatrix.jl:712-716).
"""
function mod_elements!(A::ModMatrix{T}, mod_N::Integer=-1) where {T}
    N = mod_N > 0 ? mod_N : A.N
    check(ccall((:modp_mod, LIB), Cint, (Ptr{Cvoid}, Cint, Int64, Int64, Ptr{Cvoid}, Int64, Int64),
                context().h, dtype_code(T), A.padded[1], A.padded[2], A.data.ptr, ld(A), N), "mod!")
    return A
end

"""
    change_modulus(A, new_N)

A copy of A under modulus new_N, its entries reduced mod new_N when new_N <
A.N and copied as they are otherwise (CuModMatrix.jl:723-739).
"""
function change_modulus(A::ModMatrix{T}, new_N::Integer) where {T}
    R = ModMatrix{T}(rows(A), cols(A), new_N; zero=false)
    check(ccall((:modp_memcpy_d2d, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                context().h, R.data.ptr, A.data.ptr, A.data.nbytes), "copy")
    new_N < A.N && mod_elements!(R)
    return R
end

"""
    change_modulus_no_alloc!(A, new_N)

A's storage reduced in place when new_N < A.N, returned under modulus new_N
(CuModMatrix.jl:751-758).
"""
function change_modulus_no_alloc!(A::ModMatrix{T}, new_N::Integer) where {T}
    new_N < A.N && mod_elements!(A, new_N)
    return ModMatrix{T}(A.data, A.padded, A.size, new_N)
end

# ---- GEMM (stripe_mul.jl:13-27, 175-244; CuModMatrix.jl:767-787; mul_ops.jl:54-58)
function find_max_stripe_ops(T, N)
    bits = T === Float64 ? 53 : T === Float32 ? 24 : T === Float16 ? 11 :
           throw(ArgumentError("The input type is not supported for CUBLAS gemm"))
    return floor(Int, (2^bits - 1) / (N - 1)^2) - 1
end

function gemm!(C::ModMatrix{T}, A::ModMatrix{T}, B::ModMatrix{T}, N::Integer; form=FORM_AUTO) where {T}
    check(ccall((:modp_gemm, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cint, Int64, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int64),
                context().h, form, dtype_code(T), rows(A), cols(B), cols(A),
                A.data.ptr, ld(A), B.data.ptr, ld(B), C.data.ptr, ld(C), N), "gemm")
    return C
end

function stripe_mul!(C::ModMatrix, A::ModMatrix, B::ModMatrix; M=nothing, N=nothing, form=FORM_AUTO)
    (A.N != B.N || B.N != C.N) && throw(ArgumentError("Mismatched modulus in matmul"))
    cols(A) != rows(B) && throw(DimensionMismatch(""))
    R = N === nothing ? A.N : N
    M === nothing && (M = find_max_stripe_ops(eltype(A), R))
    M < 1 && throw(ArgumentError("cannot perform a single multiplication for modulus $(A.N) with datatype $(eltype(A))"))
    return gemm!(C, A, B, R; form=form)
end

function LinearAlgebra.mul!(C::ModMatrix, A::ModMatrix, B::ModMatrix; M=nothing, N=nothing)
    (A.N != B.N || A.N != C.N) && throw(CuModArrayModulusMismatchException("All matrices must have the same modulus N"))
    cols(A) != rows(B) && throw(CuModArraySizeMismatchException("Matrix dimensions do not match for multiplication"))
    (rows(C) != rows(A) || cols(C) != cols(B)) && throw(CuModArraySizeMismatchException("Output matrix C has incorrect dimensions"))
    stripe_mul!(C, A, B; M=M, N=N)
    return C
end

function Base.:*(A::ModMatrix{T}, B::ModMatrix{T}) where {T}
    C = zeros(T, rows(A), cols(B), A.N)
    stripe_mul!(C, A, B)
    return C
end

"""
    A^n

Binary exponentiation as in CuModMatrix.jl:307-329: a square matrix only;
A^0 is the identity, a negative power is a power of the inverse.
"""
function Base.:^(A::ModMatrix{T}, n::Integer) where {T}
    rows(A) != cols(A) && throw(CuModMatrixNotSquareException("Matrix must be square for power operation"))
    if n == 0
        return eye(T, rows(A), A.N)
    elseif n < 0
        return inverse(A)^(-n)
    elseif n == 1
        return A
    elseif n % 2 == 0
        H = A^(n ÷ 2)
        return H * H
    else
        return A * (A^(n - 1))
    end
end

function mat_mul_gpu_type(A::ModMatrix{T}, B::ModMatrix{T}, mod_N::Integer=-1) where {T}
    N = mod_N > 0 ? mod_N : A.N
    C = zeros(T, rows(A), cols(B), N)
    return gemm!(C, A, B, N)
end

mat_mul_type_inplace!(C::ModMatrix, A::ModMatrix, B::ModMatrix, mod_N::Integer=-1) =
    gemm!(C, A, B, mod_N > 0 ? mod_N : A.N)

# ---- elimination (rref_lu_pluq/pluq_kernels.jl:11-157) -------------------------
function mod_inv(p, P)
    inv, new_inv, rem, new_rem = 0, 1, P, p
    while new_rem != 0
        q = div(rem, new_rem)
        inv, new_inv = new_inv, inv - q * new_inv
        rem, new_rem = new_rem, rem - q * new_rem
    end
    return inv < 0 ? inv + P : inv
end

swap_stack(buf, n) = [(Int(buf[2i-1]), Int(buf[2i])) for i in 1:n]

function pluq(A::ModMatrix{T}; form=FORM_AUTO) where {T}
    m, n = size(A)
    U = zeros(T, m, n, A.N)
    L = zeros(T, m, m, A.N)
    rs = Vector{Int64}(undef, 2 * max(1, min(m, n)))
    cs = Vector{Int64}(undef, 4 * max(1, n))
    nr, nc, rk = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:modp_pluq, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cint, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int64,
                 Ptr{Int64}, Ref{Int64}, Ptr{Int64}, Ref{Int64}, Ref{Int64}),
                context().h, form, dtype_code(T), m, n, A.data.ptr, ld(A), U.data.ptr, ld(U), L.data.ptr, ld(L), A.N,
                rs, nr, cs, nc, rk), "pluq")
    return U, L, swap_stack(rs, nr[]), swap_stack(cs, nc[]), rk[]
end

"""
    pluq_gpu_kernel(A; debug=false) -> (U, L, Perm_rows, Perm_cols)

Same outputs and conventions as the reference (pluq_kernels.jl:46-157).
"""
function pluq_gpu_kernel(A::ModMatrix; debug::Bool=false)
    U, L, P, Q, _ = pluq(A)
    return U, L, P, Q
end

function lu_ple(A::ModMatrix{T}; form=FORM_AUTO, N::Integer=A.N) where {T}
    m, n = size(A)
    U = zeros(T, m, n, A.N)
    L = zeros(T, m, m, A.N)
    perm = Vector{Int64}(undef, max(1, m))
    piv = Vector{Int64}(undef, max(1, min(m, n)))
    rk = Ref{Int64}(0)
    check(ccall((:modp_lu, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cint, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int64,
                 Ptr{Int64}, Ptr{Int64}, Ref{Int64}),
                context().h, form, dtype_code(T), m, n, A.data.ptr, ld(A), U.data.ptr, ld(U), L.data.ptr, ld(L), N,
                perm, piv, rk), "lu")
    return U, L, perm[1:m]
end

"""
    lu(A::ModMatrix) -> (U, L, perm)

PLE with row pivoting (largest residue, first row on ties), zero columns
skipped: A[perm, :] == L * U mod N.  The call surface of lu_gpu_type
(test/Experiments/rref_gpu_type.jl:60-103), which takes an optional modulus
override.
"""
LinearAlgebra.lu(A::ModMatrix; form=FORM_AUTO) = lu_ple(A; form=form)
lu_gpu_type(A::ModMatrix, mod_N::Integer=-1) = lu_ple(A; N=(mod_N > 0 ? mod_N : A.N))

function rref(A::ModMatrix{T}; form=FORM_AUTO) where {T}
    m, n = size(A)
    R = zeros(T, m, n, A.N)
    piv = Vector{Int64}(undef, max(1, min(m, n)))
    rk = Ref{Int64}(0)
    check(ccall((:modp_rref, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cint, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int64, Ptr{Int64}, Ref{Int64}),
                context().h, form, dtype_code(T), m, n, A.data.ptr, ld(A), R.data.ptr, ld(R), A.N, piv, rk), "rref")
    return R
end

# the call surface of rref_gpu_type (test/Experiments/rref_gpu_type.jl:8-51)
rref_gpu_type(A::ModMatrix) = rref(A)

# ---- permutations (permutations.jl:11-125) ------------------------------------
# ccall needs its (function, library) pair to be a constant: one method per
# C entry point, generated here
for (jlname, cname) in ((:row_perm!, :modp_apply_row_perm), (:col_perm!, :modp_apply_col_perm))
    @eval function $jlname(P::Vector{Tuple{Int,Int}}, A::ModMatrix{T}, inverse::Integer) where {T}
        st = Int64[x for t in P for x in t]
        check(ccall(($(QuoteNode(cname)), LIB), Cint,
                    (Ptr{Cvoid}, Cint, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Int64}, Int64, Cint),
                    context().h, dtype_code(T), A.padded[1], A.padded[2], A.data.ptr, ld(A), st, length(P), inverse),
              $(string(cname)))
        return nothing
    end
end

apply_row_perm!(P, A::ModMatrix) = row_perm!(P, A, 0)
apply_row_inv_perm!(P, A::ModMatrix) = row_perm!(P, A, 1)
apply_col_perm!(P, A::ModMatrix) = col_perm!(P, A, 0)
apply_col_inv_perm!(P, A::ModMatrix) = col_perm!(P, A, 1)

# ---- inverses (triangular_inverse_no_copy.jl:197-229,450-478; CuModMatrix.jl:340-502) ----
function triangular_inverse(A::ModMatrix{T}, uplo::Integer) where {T}
    m, n = size(A)
    Ainv = zeros(T, n, m, A.N)
    st = ccall((:modp_trtri, LIB), Cint,
               (Ptr{Cvoid}, Cint, Cint, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int64),
               context().h, dtype_code(T), uplo, m, n, A.data.ptr, ld(A), Ainv.data.ptr, ld(Ainv), A.N)
    st == 1 && throw(InverseNotDefinedException("(Right) pseudoinverse not defined for tall lower triangular matrices"))
    st == 4 && throw(MatrixNotInvertibleException("triangular matrix with a zero diagonal entry mod $(A.N)"))
    check(st, "triangular inverse")
    return Ainv
end

upper_triangular_inverse_no_copy(A::ModMatrix; debug::Bool=false) = triangular_inverse(A, 0)
lower_triangular_inverse_no_copy(A::ModMatrix; debug::Bool=false) = triangular_inverse(A, 1)

function inverse_or_nothing(A::ModMatrix{T}; form=FORM_AUTO) where {T}
    m, n = size(A)
    m == n || throw(CuModMatrixNotSquareException("Matrix must be square to be inverted"))
    Ainv = zeros(T, n, n, A.N)
    ok = Ref{Cint}(0)
    check(ccall((:modp_inverse, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cint, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int64, Ref{Cint}),
                context().h, form, dtype_code(T), n, A.data.ptr, ld(A), Ainv.data.ptr, ld(Ainv), A.N, ok), "inverse")
    return ok[] != 0 ? Ainv : nothing
end

function inverse(A::ModMatrix; form=FORM_AUTO)
    X = inverse_or_nothing(A; form=form)
    X === nothing && throw(MatrixNotInvertibleException("Matrix is not invertible mod $(A.N)"))
    return X
end

function is_invertible_with_inverse(A::ModMatrix; debug::Bool=false)
    X = inverse_or_nothing(A)
    return X === nothing ? (false, nothing) : (true, X)
end

is_invertible(A::ModMatrix) = pluq(A)[5] == minimum(size(A))

# ---- matrix-vector product (stripe_mul.jl:82-168, CuModMatrix.jl:816-836) ---------
# a vector is held as an n x 1 ModMatrix
function LinearAlgebra.mul!(z::ModMatrix{T}, A::ModMatrix{T}, x::ModMatrix{T}) where {T}
    (A.N == z.N && A.N == x.N) || throw(CuModArrayModulusMismatchException("All matrices must have the same modulus N"))
    (cols(A) == rows(x) && rows(z) == rows(A) && cols(x) == 1 && cols(z) == 1) ||
        throw(CuModArraySizeMismatchException("Matrix dimensions do not match for multiplication"))
    check(ccall((:modp_gemv, LIB), Cint,
                (Ptr{Cvoid}, Cint, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Int64),
                context().h, dtype_code(T), rows(A), cols(A), A.data.ptr, ld(A), x.data.ptr, z.data.ptr, A.N), "gemv")
    return z
end

# ---- Chinese remaindering of per-prime results (composite modulus, KaratsubaMatrix.jl:18-51) ----
function crt(A1::ModMatrix{T}, A2::ModMatrix{T}) where {T}
    size(A1) == size(A2) || throw(CuModArraySizeMismatchException("CRT needs matrices of the same size"))
    Y = zeros(T, rows(A1), cols(A1), A1.N * A2.N)
    check(ccall((:modp_crt2, LIB), Cint,
                (Ptr{Cvoid}, Cint, Int64, Int64, Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}, Int64),
                context().h, dtype_code(T), rows(A1), cols(A1), A1.data.ptr, ld(A1), A1.N, A2.data.ptr, ld(A2), A2.N,
                Y.data.ptr, ld(Y)), "crt")
    return Y
end

"""
    crt_reconstruct(mats; signed=false) -> Matrix{BigInt}

Chinese remaindering of residue matrices of equal size under pairwise
coprime moduli (their N, at most 12 of them): y with y == mats[i] mod N_i,
in [0, P) or, signed, in (-P/2, P/2], P = prod N_i.  The device writes the
little-endian 32-bit words of every entry (`modp_crt`); they are assembled
into BigInts here.
"""
function crt_reconstruct(mats::Vector{ModMatrix{T}}; signed::Bool=false) where {T}
    isempty(mats) && throw(ArgumentError("CRT needs at least one residue matrix"))
    m, n = size(mats[1])
    all(A -> size(A) == (m, n), mats) || throw(CuModArraySizeMismatchException("CRT needs matrices of the same size"))
    moduli = Int64[A.N for A in mats]
    words = ceil(Int, (sum(log2.(Float64.(moduli))) + (signed ? 1 : 0)) / 32 - 1e-12)
    buf = DeviceBuffer(4 * max(1, words * m * n))
    ptrs = Ptr{Cvoid}[A.data.ptr for A in mats]
    lds = Int64[ld(A) for A in mats]
    check(ccall((:modp_crt, LIB), Cint,
                (Ptr{Cvoid}, Cint, Int64, Int64, Cint, Ptr{Ptr{Cvoid}}, Ptr{Int64}, Ptr{Int64}, Cint, Ptr{Cvoid},
                 Int64, Int64),
                context().h, dtype_code(T), m, n, length(mats), ptrs, lds, moduli, signed ? 1 : 0, buf.ptr,
                max(m, 1), words), "crt")
    host = Vector{UInt32}(undef, max(1, words * m * n))
    check(ccall((:modp_memcpy_d2h, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                context().h, host, buf.ptr, sizeof(host)), "download")
    Y = Matrix{BigInt}(undef, m, n)
    top = BigInt(1) << (32 * words)
    for j in 1:n, i in 1:m
        y = BigInt(0)
        for w in words:-1:1
            y = (y << 32) + host[(w - 1) * m * n + (j - 1) * m + i]
        end
        Y[i, j] = (signed && y >= top >> 1) ? y - top : y
    end
    return Y
end

end # module
