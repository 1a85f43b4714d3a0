// util.hip -- elementwise kernels: mod, seeded rand, eye, storage conversion,
// transposition-stack permutations, two-modulus Chinese remaindering.
#include <algorithm>
#include <vector>

#include "common.h"

namespace {

constexpr int kThreads = 256;

template <typename T> __device__ __forceinline__ T *at(void *X, int64_t ld, int64_t i, int64_t j)
{
    return reinterpret_cast<T *>(X) + i + j * ld;
}

// X = mod(X, N), floor semantics (src/CuModMatrix/kernel_ops/mod_ops.jl:3-11).
// Grid-stride over the rows x cols block; one column segment per thread row.
template <typename T>
__global__ void k_mod(int64_t rows, int64_t cols, T *X, int64_t ldx, ModP mp)
{
    int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = e % rows, j = e / rows;
        T &x = X[i + j * ldx];
        x = (T)dmod((double)x, mp);
    }
}

// counter-based SplitMix64 + Lemire unbiased draw in [0,p)
__device__ __forceinline__ uint64_t sm_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t draw(uint64_t seed, uint64_t index, uint64_t p)
{
    const uint64_t gamma = 0x9E3779B97F4A7C15ULL;
    uint64_t x = sm_mix(seed + (index + 1) * gamma);
    uint64_t lo = x * p, hi = __umul64hi(x, p);
    if (lo < p) {
        uint64_t t = (0 - p) % p;
        while (lo < t) {
            x = sm_mix(x + gamma);
            lo = x * p;
            hi = __umul64hi(x, p);
        }
    }
    return hi;
}

template <typename T>
__global__ void k_rand(int64_t rows, int64_t cols, T *X, int64_t ldx, uint64_t p, uint64_t seed,
                       uint64_t offset)
{
    int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = e % rows, j = e / rows;
        X[i + j * ldx] = (T)draw(seed, offset + (uint64_t)e, p);
    }
}

template <typename T>
__global__ void k_eye(int64_t rows, int64_t cols, T *X, int64_t ldx)
{
    int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = e % rows, j = e / rows;
        X[i + j * ldx] = (T)(i == j ? 1 : 0);
    }
}

template <typename TI, typename TO>
__global__ void k_convert(int64_t rows, int64_t cols, const TI *X, int64_t ldx, TO *Y, int64_t ldy,
                          ModP mp)
{
    int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = e % rows, j = e / rows;
        Y[i + j * ldy] = from_res<TO>(to_res<TI>(X[i + j * ldx], mp));
    }
}

// Row permutation by a transposition stack: one thread per column, the
// stack applied in order (or reversed), as permutations.jl:112-125.
template <typename T>
__global__ void k_perm_rows(int64_t rows, int64_t cols, T *X, int64_t ldx, const int64_t *stack,
                            int64_t nswaps, int inverse)
{
    int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= cols) return;
    T *c = X + j * ldx;
    for (int64_t s = 0; s < nswaps; ++s) {
        int64_t e = inverse ? nswaps - 1 - s : s;
        int64_t a = stack[2 * e] - 1, b = stack[2 * e + 1] - 1;
        T t = c[a]; c[a] = c[b]; c[b] = t;
    }
}

// Column permutation: one thread per row (permutations.jl:49-62).
template <typename T>
__global__ void k_perm_cols(int64_t rows, int64_t cols, T *X, int64_t ldx, const int64_t *stack,
                            int64_t nswaps, int inverse)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= rows) return;
    for (int64_t s = 0; s < nswaps; ++s) {
        int64_t e = inverse ? nswaps - 1 - s : s;
        int64_t a = stack[2 * e] - 1, b = stack[2 * e + 1] - 1;
        T t = X[i + a * ldx]; X[i + a * ldx] = X[i + b * ldx]; X[i + b * ldx] = t;
    }
}

// Lines (columns, or rows when rowwise) moved by a permutation, through a
// compact buffer: first every source line is copied out, then written to its
// destination, so the moves may overlap in any order.  pairs[2k] is the
// destination line, pairs[2k+1] the source; len elements per line.
template <typename T>
__global__ void k_lines_out(int rowwise, int64_t len, const T *__restrict__ X, int64_t ldx,
                            const int64_t *__restrict__ pairs, int64_t nm, T *__restrict__ tmp)
{
    const int64_t total = nm * len;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = e / len, t = e % len, src = pairs[2 * k + 1];
        tmp[e] = rowwise ? X[src + t * ldx] : X[t + src * ldx];
    }
}
template <typename T>
__global__ void k_lines_in(int rowwise, int64_t len, T *__restrict__ X, int64_t ldx,
                           const int64_t *__restrict__ pairs, int64_t nm, const T *__restrict__ tmp)
{
    const int64_t total = nm * len;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = e / len, t = e % len, dst = pairs[2 * k];
        (rowwise ? X[dst + t * ldx] : X[t + dst * ldx]) = tmp[e];
    }
}

unsigned grid_for(int64_t total)
{
    int64_t b = (total + kThreads - 1) / kThreads;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (unsigned)b;
}

}  // namespace

int util_mod(modp_ctx *ctx, int kind, int64_t rows, int64_t cols, void *X, int64_t ldx, int64_t N)
{
    if (rows <= 0 || cols <= 0) return MODP_OK;
    ModP mp = make_modp(N);
    unsigned g = grid_for(rows * cols);
    if (kind == EK_F64)
        k_mod<double><<<g, kThreads, 0, ctx->stream>>>(rows, cols, (double *)X, ldx, mp);
    else if (kind == EK_F32)
        k_mod<float><<<g, kThreads, 0, ctx->stream>>>(rows, cols, (float *)X, ldx, mp);
    else
        return set_error(ctx, MODP_ERR_DTYPE, "mod: unsupported element type");
    HIP_TRY(ctx, hipGetLastError());
    return MODP_OK;
}

int util_rand(modp_ctx *ctx, int kind, int64_t rows, int64_t cols, void *X, int64_t ldx, int64_t p,
              uint64_t seed, uint64_t offset)
{
    if (rows <= 0 || cols <= 0) return MODP_OK;
    unsigned g = grid_for(rows * cols);
    if (kind == EK_F64)
        k_rand<double><<<g, kThreads, 0, ctx->stream>>>(rows, cols, (double *)X, ldx, (uint64_t)p, seed, offset);
    else if (kind == EK_F32)
        k_rand<float><<<g, kThreads, 0, ctx->stream>>>(rows, cols, (float *)X, ldx, (uint64_t)p, seed, offset);
    else if (kind == EK_I32)
        k_rand<int32_t><<<g, kThreads, 0, ctx->stream>>>(rows, cols, (int32_t *)X, ldx, (uint64_t)p, seed, offset);
    else
        return set_error(ctx, MODP_ERR_DTYPE, "rand: unsupported element type");
    HIP_TRY(ctx, hipGetLastError());
    return MODP_OK;
}

int util_eye(modp_ctx *ctx, int kind, int64_t rows, int64_t cols, void *X, int64_t ldx)
{
    if (rows <= 0 || cols <= 0) return MODP_OK;
    unsigned g = grid_for(rows * cols);
    if (kind == EK_F64)
        k_eye<double><<<g, kThreads, 0, ctx->stream>>>(rows, cols, (double *)X, ldx);
    else if (kind == EK_F32)
        k_eye<float><<<g, kThreads, 0, ctx->stream>>>(rows, cols, (float *)X, ldx);
    else
        return set_error(ctx, MODP_ERR_DTYPE, "eye: unsupported element type");
    HIP_TRY(ctx, hipGetLastError());
    return MODP_OK;
}

// Y = the y in [0, p1 p2) with y = X1 mod p1 and y = X2 mod p2:
// y = x1 + p1 * ((x2 - x1) * c mod p2), c = p1^{-1} mod p2 (Garner's step).
template <typename T>
__global__ void k_crt2(int64_t rows, int64_t cols, const T *X1, int64_t ld1, const T *X2, int64_t ld2, T *Y,
                       int64_t ldy, ModP m1, ModP m2, double c)
{
    int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = e % rows, j = e / rows;
        const double x1 = (double)to_res<T>(X1[i + j * ld1], m1);
        const double x2 = (double)to_res<T>(X2[i + j * ld2], m2);
        const double d = dmod(x2 - dmod(x1, m2), m2);    // (x2 - x1) mod p2
        const double t = dmod(d * c, m2);                 // < p2^2 <= 2^52: exact
        Y[i + j * ldy] = (T)fma(m1.pd, t, x1);            // < p1 p2 <= 2^53: exact
    }
}

template <typename TI>
static int convert_from(modp_ctx *ctx, int64_t rows, int64_t cols, const TI *X, int64_t ldx, int kout,
                        void *Y, int64_t ldy, const ModP &mp)
{
    unsigned g = grid_for(rows * cols);
    if (kout == EK_F64)
        k_convert<TI, double><<<g, kThreads, 0, ctx->stream>>>(rows, cols, X, ldx, (double *)Y, ldy, mp);
    else if (kout == EK_F32)
        k_convert<TI, float><<<g, kThreads, 0, ctx->stream>>>(rows, cols, X, ldx, (float *)Y, ldy, mp);
    else
        k_convert<TI, int32_t><<<g, kThreads, 0, ctx->stream>>>(rows, cols, X, ldx, (int32_t *)Y, ldy, mp);
    HIP_TRY(ctx, hipGetLastError());
    return MODP_OK;
}

int util_convert(modp_ctx *ctx, int kin, int64_t rows, int64_t cols, const void *X, int64_t ldx,
                 int kout, void *Y, int64_t ldy, const ModP &mp)
{
    if (rows <= 0 || cols <= 0) return MODP_OK;
    if (kin == EK_F64) return convert_from(ctx, rows, cols, (const double *)X, ldx, kout, Y, ldy, mp);
    if (kin == EK_F32) return convert_from(ctx, rows, cols, (const float *)X, ldx, kout, Y, ldy, mp);
    return convert_from(ctx, rows, cols, (const int32_t *)X, ldx, kout, Y, ldy, mp);
}

int util_perm(modp_ctx *ctx, int kind, int rowwise, int64_t rows, int64_t cols, void *X, int64_t ldx,
              const int64_t *stack_host, int64_t nswaps, int inverse)
{
    if (nswaps <= 0 || rows <= 0 || cols <= 0) return MODP_OK;
    int64_t lim = rowwise ? rows : cols;
    for (int64_t s = 0; s < 2 * nswaps; ++s)
        if (stack_host[s] < 1 || stack_host[s] > lim)
            return set_error(ctx, MODP_ERR_ARG, "permutation index %lld out of range 1..%lld",
                             (long long)stack_host[s], (long long)lim);
    int st = MODP_OK;
    if (nswaps > 8) {
        // many swaps: compose them on the host (the line that ends at j came
        // from idx[j]) and move each displaced line once
        std::vector<int64_t> idx((size_t)lim);
        for (int64_t i = 0; i < lim; ++i) idx[(size_t)i] = i;
        for (int64_t s = 0; s < nswaps; ++s) {
            const int64_t e = inverse ? nswaps - 1 - s : s;
            std::swap(idx[(size_t)(stack_host[2 * e] - 1)], idx[(size_t)(stack_host[2 * e + 1] - 1)]);
        }
        std::vector<int64_t> pairs;
        for (int64_t j = 0; j < lim; ++j)
            if (idx[(size_t)j] != j) {
                pairs.push_back(j);
                pairs.push_back(idx[(size_t)j]);
            }
        const int64_t nm = (int64_t)pairs.size() / 2;
        if (nm == 0) return MODP_OK;
        const int64_t len = rowwise ? cols : rows;
        const size_t es = kind == EK_F64 ? 8 : 4;
        int64_t *dpairs = (int64_t *)ws_get(ctx, WS_MISC, sizeof(int64_t) * 2 * (size_t)nm, &st);
        if (!dpairs) return st;
        void *tmp = ws_get(ctx, WS_PERM, es * (size_t)nm * (size_t)len, &st);
        if (!tmp) return st;
        HIP_TRY(ctx, hipMemcpyAsync(dpairs, pairs.data(), sizeof(int64_t) * 2 * (size_t)nm, hipMemcpyHostToDevice,
                                    ctx->stream));
        const unsigned g = grid_for(nm * len);
#define LINES(T)                                                                                              \
    k_lines_out<T><<<g, kThreads, 0, ctx->stream>>>(rowwise, len, (const T *)X, ldx, dpairs, nm, (T *)tmp); \
    k_lines_in<T><<<g, kThreads, 0, ctx->stream>>>(rowwise, len, (T *)X, ldx, dpairs, nm, (const T *)tmp);
        if (kind == EK_F64) {
            LINES(double)
        } else if (kind == EK_F32) {
            LINES(float)
        } else {
            LINES(int32_t)
        }
#undef LINES
        HIP_TRY(ctx, hipGetLastError());
        // the pair buffer is reused by the next call and the host vector
        // must outlive the copy: keep the host program order
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        return MODP_OK;
    }
    int64_t *dstack = (int64_t *)ws_get(ctx, WS_MISC, sizeof(int64_t) * 2 * nswaps, &st);
    if (!dstack) return st;
    HIP_TRY(ctx, hipMemcpyAsync(dstack, stack_host, sizeof(int64_t) * 2 * nswaps, hipMemcpyHostToDevice,
                                ctx->stream));
    int64_t work = rowwise ? cols : rows;
    unsigned g = (unsigned)((work + kThreads - 1) / kThreads);
    if (kind == EK_F64) {
        if (rowwise) k_perm_rows<double><<<g, kThreads, 0, ctx->stream>>>(rows, cols, (double *)X, ldx, dstack, nswaps, inverse);
        else k_perm_cols<double><<<g, kThreads, 0, ctx->stream>>>(rows, cols, (double *)X, ldx, dstack, nswaps, inverse);
    } else if (kind == EK_F32) {
        if (rowwise) k_perm_rows<float><<<g, kThreads, 0, ctx->stream>>>(rows, cols, (float *)X, ldx, dstack, nswaps, inverse);
        else k_perm_cols<float><<<g, kThreads, 0, ctx->stream>>>(rows, cols, (float *)X, ldx, dstack, nswaps, inverse);
    } else {
        if (rowwise) k_perm_rows<int32_t><<<g, kThreads, 0, ctx->stream>>>(rows, cols, (int32_t *)X, ldx, dstack, nswaps, inverse);
        else k_perm_cols<int32_t><<<g, kThreads, 0, ctx->stream>>>(rows, cols, (int32_t *)X, ldx, dstack, nswaps, inverse);
    }
    HIP_TRY(ctx, hipGetLastError());
    // the stack buffer is reused by the next call: keep the host program
    // order by waiting for this one
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MODP_OK;
}

int util_crt2(modp_ctx *ctx, int kind, int64_t rows, int64_t cols, const void *X1, int64_t ld1, int64_t p1,
              const void *X2, int64_t ld2, int64_t p2, void *Y, int64_t ldy)
{
    // c = p1^{-1} mod p2 by the extended Euclidean algorithm on the host
    int64_t r0 = p2, r1 = p1 % p2, s0 = 0, s1 = 1;
    while (r1 != 0) {
        const int64_t q = r0 / r1;
        int64_t t = r0 - q * r1; r0 = r1; r1 = t;
        t = s0 - q * s1; s0 = s1; s1 = t;
    }
    if (r0 != 1)
        return set_error(ctx, MODP_ERR_ARG, "moduli %lld and %lld are not coprime", (long long)p1, (long long)p2);
    const int64_t c = ((s0 % p2) + p2) % p2;
    if (rows <= 0 || cols <= 0) return MODP_OK;
    const unsigned g = grid_for(rows * cols);
    const ModP m1 = make_modp(p1), m2 = make_modp(p2);
    if (kind == EK_F64)
        k_crt2<double><<<g, kThreads, 0, ctx->stream>>>(rows, cols, (const double *)X1, ld1, (const double *)X2, ld2,
                                                         (double *)Y, ldy, m1, m2, (double)c);
    else
        k_crt2<float><<<g, kThreads, 0, ctx->stream>>>(rows, cols, (const float *)X1, ld1, (const float *)X2, ld2,
                                                        (float *)Y, ldy, m1, m2, (double)c);
    return hip_check(ctx, hipGetLastError(), "k_crt2 launch");
}
