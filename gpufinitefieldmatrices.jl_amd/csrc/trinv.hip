// trinv.hip -- inverses mod p: triangular matrices (64-row diagonal blocks
// inverted in LDS, then the off-diagonal blocks by GEMM, level by level with
// the independent products of a level on side streams) and square matrices
// through PLUQ, A^{-1} = Q U^{-1} L^{-1} P.
//
// Reference: upper/lower_triangular_inverse_no_copy
// (src/CuModMatrix/triangular/triangular_inverse_no_copy.jl:18-478), whose
// recursion splits at a multiple of its tile width and combines with
// X12 = -X11 T12 X22; and inverse / is_invertible_with_inverse
// (src/CuModMatrix/CuModMatrix.jl:351-502), which multiplies the permuted
// triangular inverses.  Inverses are unique, so the split and the GEMM form
// do not change the result.
#include <algorithm>
#include <vector>

#include "common.h"

namespace {

constexpr int TB = 64;  // diagonal block edge

// Inverse of each TB x TB diagonal block of the triangular int32 matrix T
// (k x k, leading dimension ldt) into the same block of X.  One workgroup
// per block; thread c solves column c of the block inverse by substitution,
// keeping the column in LDS.  A zero diagonal entry sets *bad.
template <bool UPPER>
__global__ __launch_bounds__(TB) void k_trinv_diag(const int32_t *__restrict__ T, int64_t ldt,
                                                   int32_t *__restrict__ X, int64_t ldx, int64_t k, ModP mp,
                                                   int *__restrict__ bad)
{
    __shared__ uint32_t S[TB][TB + 1];   // S[r][c] = T(o + r, o + c)
    __shared__ uint32_t Xs[TB][TB + 1];  // Xs[r][c] = X(o + r, o + c)
    __shared__ uint32_t dinv[TB];
    const int c = threadIdx.x;
    const int64_t o = (int64_t)blockIdx.x * TB;
    const int nb = (int)(k - o < TB ? k - o : TB);
    // sixteen loads in flight at a time (indices clamped into the block, so
    // no load waits behind a branch)
    for (int j0 = 0; j0 < nb; j0 += 16) {
        int32_t v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = T[(o + min(c, nb - 1)) + (o + min(j0 + u, nb - 1)) * ldt];
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (j0 + u < nb) S[c][j0 + u] = c < nb ? (uint32_t)v[u] : 0u;
    }
    __syncthreads();
    if (c < nb) {
        const uint32_t d = S[c][c];
        const uint32_t di = d ? modinv(d, mp) : 0u;
        dinv[c] = di;
        if (!di) *bad = 1;
    }
    __syncthreads();
    if (c < nb) {
        if (UPPER) {
            // rows c, c-1, ..., 0 of column c; rows below c are zero
            for (int r = nb - 1; r > c; --r) Xs[r][c] = 0u;
            for (int r = c; r >= 0; --r) {
                uint32_t s = r == c ? 1u : 0u;
                for (int q = r + 1; q <= c; ++q) s = submod(s, mulmod(S[r][q], Xs[q][c], mp), mp);
                Xs[r][c] = mulmod(s, dinv[r], mp);
            }
        } else {
            // rows c, c+1, ..., nb-1 of column c; rows above c are zero
            for (int r = 0; r < c; ++r) Xs[r][c] = 0u;
            for (int r = c; r < nb; ++r) {
                uint32_t s = r == c ? 1u : 0u;
                for (int q = c; q < r; ++q) s = submod(s, mulmod(S[r][q], Xs[q][c], mp), mp);
                Xs[r][c] = mulmod(s, dinv[r], mp);
            }
        }
    }
    __syncthreads();
    for (int j = 0; j < nb; ++j)
        if (c < nb) X[(o + c) + (o + j) * ldx] = (int32_t)Xs[c][j];
}

int64_t rup64(int64_t x) { return (x + 63) / 64 * 64; }

GemmArgs i32_gemm(int64_t m, int64_t n, int64_t k, const int32_t *A, int64_t lda, const int32_t *B, int64_t ldb,
                  int32_t *C, int64_t ldc, int epi, const ModP &mp)
{
    GemmArgs g;
    g.m = m; g.n = n; g.k = k;
    g.A = A; g.lda = lda; g.ka = EK_I32;
    g.B = B; g.ldb = ldb; g.kb = EK_I32;
    g.C = C; g.ldc = ldc; g.kc = EK_I32;
    g.epi = epi;
    g.mp = mp;
    return g;
}

// Off-diagonal blocks of X = T^{-1}, bottom-up by levels, after the diagonal
// blocks are in place and X is zero elsewhere.  At level b the blocks
// [lo, lo + 2b) (lo a multiple of 2b) are completed from their halves:
// upper X12 = -X11 (T12 X22), lower X21 = -X22 (T21 X11), with X11 and X22
// finished at the level below.  The pairs of one level are independent:
// while they are small (b < 512) they run on the side streams, each pair's
// two products on one stream in the plain form without K slices (no shared
// workspace), with a fork and a join per level; larger levels run on the
// library stream with the automatic form.
int set_stream(modp_ctx *ctx, hipStream_t s)
{
    ctx->stream = s;
    return MODP_OK;
}

int side_streams(modp_ctx *ctx)
{
    if (ctx->fork_ev) return MODP_OK;
    for (int i = 0; i < modp_ctx::NSIDE; ++i) {
        HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->side[i], hipStreamNonBlocking));
        HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->join_ev[i], hipEventDisableTiming));
    }
    HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
    return MODP_OK;
}

int trinv_pair(modp_ctx *ctx, bool upper, const int32_t *T, int64_t ldt, int32_t *X, int64_t ldx, int32_t *tmp,
               int64_t lo, int64_t h, int64_t r, const ModP &mp, bool plain)
{
    const int32_t *T11 = T + lo + lo * ldt;
    const int32_t *X11 = X + lo + lo * ldx, *X22 = X + (lo + h) + (lo + h) * ldx;
    GemmArgs g1, g2;
    if (upper) {
        // tmp (h x r) = T12 X22; X12 = 0 - X11 tmp
        g1 = i32_gemm(h, r, r, T11 + h * ldt, ldt, X22, ldx, tmp, h, EPI_STORE, mp);
        g2 = i32_gemm(h, r, h, X11, ldx, tmp, h, X + lo + (lo + h) * ldx, ldx, EPI_SUB, mp);
    } else {
        // tmp (r x h) = T21 X11; X21 = 0 - X22 tmp
        g1 = i32_gemm(r, h, h, T11 + h, ldt, X11, ldx, tmp, r, EPI_STORE, mp);
        g2 = i32_gemm(r, h, r, X22, ldx, tmp, r, X + (lo + h) + lo * ldx, ldx, EPI_SUB, mp);
    }
    for (const GemmArgs *g : {&g1, &g2}) {
        GemmConfig cfg = gemm_default_config(ctx, *g, plain ? MODP_FORM_PLAIN_F64 : MODP_FORM_AUTO);
        if (plain) cfg.ksplit = 1;
        MODP_TRY(gemm_run(ctx, *g, cfg));
    }
    return MODP_OK;
}

int trinv_levels(modp_ctx *ctx, bool upper, const int32_t *T, int64_t ldt, int32_t *X, int64_t ldx, int32_t *tmp,
                 int64_t k, const ModP &mp)
{
    MODP_TRY(side_streams(ctx));
    hipStream_t main_stream = ctx->stream;
    for (int64_t b = TB; b < k; b *= 2) {
        std::vector<int64_t> los;
        for (int64_t lo = 0; lo + b < k; lo += 2 * b) los.push_back(lo);
        const bool par = b < 512 && los.size() > 1;
        if (!par) {
            for (size_t i = 0; i < los.size(); ++i)
                MODP_TRY(trinv_pair(ctx, upper, T, ldt, X, ldx, tmp, los[i], b, std::min(b, k - los[i] - b), mp, false));
            continue;
        }
        HIP_TRY(ctx, hipEventRecord(ctx->fork_ev, main_stream));
        for (int i = 0; i < modp_ctx::NSIDE; ++i) HIP_TRY(ctx, hipStreamWaitEvent(ctx->side[i], ctx->fork_ev, 0));
        int st = MODP_OK;
        for (size_t i = 0; i < los.size() && st == MODP_OK; ++i) {
            set_stream(ctx, ctx->side[i % modp_ctx::NSIDE]);
            // each pair has its own b x b slice of the temporary
            st = trinv_pair(ctx, upper, T, ldt, X, ldx, tmp + (int64_t)i * b * b, los[i], b,
                            std::min(b, k - los[i] - b), mp, true);
        }
        set_stream(ctx, main_stream);
        for (int i = 0; i < modp_ctx::NSIDE; ++i) {
            HIP_TRY(ctx, hipEventRecord(ctx->join_ev[i], ctx->side[i]));
            HIP_TRY(ctx, hipStreamWaitEvent(main_stream, ctx->join_ev[i], 0));
        }
        MODP_TRY(st);
    }
    return MODP_OK;
}

int zero2d(modp_ctx *ctx, void *X, int64_t ldx, size_t es, int64_t rows, int64_t cols)
{
    if (rows <= 0 || cols <= 0) return MODP_OK;
    return hip_check(ctx, hipMemset2DAsync(X, (size_t)ldx * es, 0, (size_t)rows * es, (size_t)cols, ctx->stream),
                     "hipMemset2DAsync");
}

size_t elem_size(int kind) { return kind == EK_F64 ? 8 : 4; }

}  // namespace

// X (k x k int32, leading dimension ldx) = inverse of the triangular k x k
// int32 matrix T; entries of T outside the triangle are ignored.
static int trinv_i32(modp_ctx *ctx, bool upper, const int32_t *T, int64_t ldt, int32_t *X, int64_t ldx, int64_t k,
                     const ModP &mp)
{
    int st = MODP_OK;
    int *bad = (int *)ws_get(ctx, WS_TRI_FLAG, 64, &st);
    if (!bad) return st;
    int32_t *tmp = (int32_t *)ws_get(ctx, WS_TRI_TMP, sizeof(int32_t) * (size_t)((k + 1) / 2 + TB) * (size_t)(k + TB), &st);
    if (!tmp) return st;
    MODP_TRY(zero2d(ctx, X, ldx, 4, k, k));
    HIP_TRY(ctx, hipMemsetAsync(bad, 0, sizeof(int), ctx->stream));
    const unsigned nbk = (unsigned)((k + TB - 1) / TB);
    if (upper) k_trinv_diag<true><<<nbk, TB, 0, ctx->stream>>>(T, ldt, X, ldx, k, mp, bad);
    else k_trinv_diag<false><<<nbk, TB, 0, ctx->stream>>>(T, ldt, X, ldx, k, mp, bad);
    HIP_TRY(ctx, hipGetLastError());
    int hbad = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (hbad)
        return set_error(ctx, MODP_ERR_NOT_INVERTIBLE, "triangular matrix has a zero (or non-invertible) diagonal entry modulo %lld",
                         (long long)mp.p);
    return trinv_levels(ctx, upper, T, ldt, X, ldx, tmp, k, mp);
}

int tri_inverse(modp_ctx *ctx, int kind, int upper, int64_t m, int64_t n, const void *A, int64_t lda, void *Ainv,
                int64_t ldai, int64_t p)
{
    if (!upper && m > n)
        return set_error(ctx, MODP_ERR_DIM, "(Right) pseudoinverse not defined for tall lower triangular matrices");
    const int64_t k = m < n ? m : n;
    // Ainv is n x m: the leading k x k block holds the inverse, zeros elsewhere
    MODP_TRY(zero2d(ctx, Ainv, ldai, elem_size(kind), n, m));
    if (k == 0) return MODP_OK;
    const ModP mp = make_modp(p);
    const int64_t ld = rup64(k);
    int st = MODP_OK;
    int32_t *T = (int32_t *)ws_get(ctx, WS_TRI_T, sizeof(int32_t) * (size_t)ld * (size_t)k, &st);
    if (!T) return st;
    int32_t *X = (int32_t *)ws_get(ctx, WS_TRI_X, sizeof(int32_t) * (size_t)ld * (size_t)k, &st);
    if (!X) return st;
    MODP_TRY(util_convert(ctx, kind, k, k, A, lda, EK_I32, T, ld, mp));
    MODP_TRY(trinv_i32(ctx, upper != 0, T, ld, X, ld, k, mp));
    return util_convert(ctx, EK_I32, k, k, X, ld, kind, Ainv, ldai, mp);
}

int mat_inverse(modp_ctx *ctx, int form, int kind, int64_t n, const void *A, int64_t lda, void *Ainv, int64_t ldai,
                int64_t p, int *invertible)
{
    *invertible = 0;
    if (n == 0) {
        *invertible = 1;
        return MODP_OK;
    }
    const ModP mp = make_modp(p);
    const int64_t ld = rup64(n);
    const size_t es = elem_size(kind);
    int st = MODP_OK;
    void *U = ws_get(ctx, WS_INV_U, es * (size_t)ld * (size_t)n, &st);
    if (!U) return st;
    void *L = ws_get(ctx, WS_INV_L, es * (size_t)ld * (size_t)n, &st);
    if (!L) return st;
    MODP_TRY(zero2d(ctx, U, ld, es, n, n));
    MODP_TRY(zero2d(ctx, L, ld, es, n, n));
    ElimResult r;
    MODP_TRY(elim_run(ctx, ELIM_PLUQ, form, kind, n, n, A, lda, U, ld, L, ld, p, r));
    if (r.rank < n) return MODP_OK;
    *invertible = 1;
    // U^{-1} and L^{-1} as int32, in the triangular workspaces' places
    int32_t *T = (int32_t *)ws_get(ctx, WS_TRI_T, sizeof(int32_t) * (size_t)ld * (size_t)n, &st);
    if (!T) return st;
    int32_t *Xu = (int32_t *)ws_get(ctx, WS_TRI_X, sizeof(int32_t) * (size_t)ld * (size_t)n, &st);
    if (!Xu) return st;
    int32_t *Xl = (int32_t *)ws_get(ctx, WS_TRI_X2, sizeof(int32_t) * (size_t)ld * (size_t)n, &st);
    if (!Xl) return st;
    MODP_TRY(util_convert(ctx, kind, n, n, U, ld, EK_I32, T, ld, mp));
    MODP_TRY(trinv_i32(ctx, true, T, ld, Xu, ld, n, mp));
    MODP_TRY(util_convert(ctx, kind, n, n, L, ld, EK_I32, T, ld, mp));
    MODP_TRY(trinv_i32(ctx, false, T, ld, Xl, ld, n, mp));
    // A^{-1} = Q U^{-1} L^{-1} P: apply_row_inv_perm!(col stack, U^{-1}) and
    // apply_col_inv_perm!(row stack, L^{-1}), then the product
    MODP_TRY(util_perm(ctx, EK_I32, 1, n, n, Xu, ld, r.col_swaps.data(), (int64_t)r.col_swaps.size() / 2, 1));
    MODP_TRY(util_perm(ctx, EK_I32, 0, n, n, Xl, ld, r.row_swaps.data(), (int64_t)r.row_swaps.size() / 2, 1));
    GemmArgs g;
    g.m = n; g.n = n; g.k = n;
    g.A = Xu; g.lda = ld; g.ka = EK_I32;
    g.B = Xl; g.ldb = ld; g.kb = EK_I32;
    g.C = Ainv; g.ldc = ldai; g.kc = kind;
    g.epi = EPI_STORE;
    g.mp = mp;
    GemmConfig cfg = gemm_default_config(ctx, g, form == MODP_FORM_AUTO ? MODP_FORM_AUTO : form);
    return gemm_run(ctx, g, cfg);
}
