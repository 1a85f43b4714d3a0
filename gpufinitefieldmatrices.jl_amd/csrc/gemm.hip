// gemm.hip -- mod-p GEMM cores for gfx950.
//
// Plain form: FP64 MFMA (v_mfma_f64_16x16x4f64) on balanced residues in
// [-floor(p/2), floor(p/2)], with the accumulators reduced every R k4-steps
// so that every partial sum stays an exact integer below 2^53 (the
// reference's stripe bound, src/CuModMatrix/kernel_mul/stripe_mul.jl:13-27,
// applied per MFMA step instead of per cuBLAS stripe).
//
// Limb form (kernels in gemm_limb.h): residues as signed base-256 digits in
// int8 on v_mfma_i32_16x16x64_i8; this file splits the operands into digit
// planes, cuts K into int32-safe chunks and picks the launch.
//
// Both forms can cut K into slices that run concurrently (blockIdx.y) and are
// summed mod p afterwards, for products whose output grid alone leaves CUs
// idle.
#include "gemm_limb.h"
#include "gemm_tile.h"

namespace {

// --------------------------------------------------------------------------
// plain FP64 form
// --------------------------------------------------------------------------

// blockIdx.y selects a K slice [y*kslice, (y+1)*kslice); slice y writes its
// residues at C + y*cslice (one slice: kslice >= k, cslice = 0).
template <typename TA, typename TB, typename TC, int BM, int BN, int EPI>
__global__ __launch_bounds__(256) void k_gemm_f64(int64_t m, int64_t n, int64_t k,
                                                  const TA *__restrict__ A, int64_t lda,
                                                  const TB *__restrict__ B, int64_t ldb,
                                                  TC *__restrict__ C, int64_t ldc, ModP mp, int R,
                                                  int64_t kslice, int64_t cslice,
                                                  const int *__restrict__ skip)
{
    if (skip && *skip) return;
    int64_t tm, tn;
    if (!tile_of_block((m + BM - 1) / BM, (n + BN - 1) / BN, &tm, &tn)) return;
    const int64_t kb = (int64_t)blockIdx.y * kslice;
    const int64_t klen = k - kb < kslice ? k - kb : kslice;
    __shared__ F64Smem<BM, BN> sm;
    gemm_f64_tile<TA, TB, TC, BM, BN, EPI>(m, n, klen, A + kb * lda, lda, B + kb, ldb, C + blockIdx.y * cslice, ldc,
                                           mp, R, tm * BM, tn * BN, sm);
}

// C = epi(C, sum of S slice results mod p); slice s holds m x n residues at
// P + s*m*n (leading dimension m).
template <typename TC, int EPI>
__global__ __launch_bounds__(256) void k_slice_sum(int64_t m, int64_t n, const int32_t *__restrict__ P, int S,
                                                   TC *__restrict__ C, int64_t ldc, ModP mp,
                                                   const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t mn = m * n;
    if (idx >= mn) return;
    double acc = 0.0;  // S * p < 2^53
    for (int s = 0; s < S; ++s) acc += (double)P[(int64_t)s * mn + idx];
    uint32_t r = (uint32_t)dmod(acc, mp);
    TC *dst = C + idx % m + (idx / m) * ldc;
    if (EPI == EPI_SUB) r = submod(to_res<TC>(*dst, mp), r, mp);
    else if (EPI == EPI_ADD) r = addmod(to_res<TC>(*dst, mp), r, mp);
    *dst = from_res<TC>(r);
}

int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// ---- dispatch helpers -----------------------------------------------------

template <typename TA, typename TB, typename TC, int BM, int BN>
int launch_f64_t(modp_ctx *ctx, const GemmArgs &g, int R, const Slices &sl)
{
    const dim3 grid(xcd_grid(((g.n + BN - 1) / BN) * ((g.m + BM - 1) / BM)), (unsigned)sl.S);
    TimedLaunch tl(ctx, "gemm_f64");
    const TA *A = (const TA *)g.A;
    const TB *B = (const TB *)g.B;
    TC *C = (TC *)g.C;
    switch (g.epi) {
    case EPI_STORE:
        k_gemm_f64<TA, TB, TC, BM, BN, EPI_STORE><<<grid, 256, 0, ctx->stream>>>(
            g.m, g.n, g.k, A, g.lda, B, g.ldb, C, g.ldc, g.mp, R, sl.kslice, sl.cslice, g.skip);
        break;
    case EPI_SUB:
        k_gemm_f64<TA, TB, TC, BM, BN, EPI_SUB><<<grid, 256, 0, ctx->stream>>>(
            g.m, g.n, g.k, A, g.lda, B, g.ldb, C, g.ldc, g.mp, R, sl.kslice, sl.cslice, g.skip);
        break;
    default:
        k_gemm_f64<TA, TB, TC, BM, BN, EPI_ADD><<<grid, 256, 0, ctx->stream>>>(
            g.m, g.n, g.k, A, g.lda, B, g.ldb, C, g.ldc, g.mp, R, sl.kslice, sl.cslice, g.skip);
        break;
    }
    return hip_check(ctx, hipGetLastError(), "k_gemm_f64 launch");
}

template <typename TA, typename TB, typename TC>
int launch_f64_tile(modp_ctx *ctx, const GemmArgs &g, int R, int tile, const Slices &sl)
{
    if (tile == 128) return launch_f64_t<TA, TB, TC, 128, 128>(ctx, g, R, sl);
    return launch_f64_t<TA, TB, TC, 64, 64>(ctx, g, R, sl);
}

template <typename TA, typename TB>
int launch_f64_c(modp_ctx *ctx, const GemmArgs &g, int R, int tile, const Slices &sl)
{
    if (g.kc == EK_F64) return launch_f64_tile<TA, TB, double>(ctx, g, R, tile, sl);
    if (g.kc == EK_F32) return launch_f64_tile<TA, TB, float>(ctx, g, R, tile, sl);
    return launch_f64_tile<TA, TB, int32_t>(ctx, g, R, tile, sl);
}

template <typename TA>
int launch_f64_b(modp_ctx *ctx, const GemmArgs &g, int R, int tile, const Slices &sl)
{
    if (g.kb == EK_F64) return launch_f64_c<TA, double>(ctx, g, R, tile, sl);
    if (g.kb == EK_F32) return launch_f64_c<TA, float>(ctx, g, R, tile, sl);
    return launch_f64_c<TA, int32_t>(ctx, g, R, tile, sl);
}

int launch_f64(modp_ctx *ctx, const GemmArgs &g, int R, int tile, const Slices &sl)
{
    if (g.ka == EK_F64) return launch_f64_b<double>(ctx, g, R, tile, sl);
    if (g.ka == EK_F32) return launch_f64_b<float>(ctx, g, R, tile, sl);
    return launch_f64_b<int32_t>(ctx, g, R, tile, sl);
}

template <typename TC>
int launch_slice_sum_t(modp_ctx *ctx, const GemmArgs &g, int epi, const int32_t *P, int S)
{
    const int64_t mn = g.m * g.n;
    const unsigned grid = (unsigned)((mn + 255) / 256);
    TC *C = (TC *)g.C;
    if (epi == EPI_STORE)
        k_slice_sum<TC, EPI_STORE><<<grid, 256, 0, ctx->stream>>>(g.m, g.n, P, S, C, g.ldc, g.mp, g.skip);
    else if (epi == EPI_SUB)
        k_slice_sum<TC, EPI_SUB><<<grid, 256, 0, ctx->stream>>>(g.m, g.n, P, S, C, g.ldc, g.mp, g.skip);
    else
        k_slice_sum<TC, EPI_ADD><<<grid, 256, 0, ctx->stream>>>(g.m, g.n, P, S, C, g.ldc, g.mp, g.skip);
    return hip_check(ctx, hipGetLastError(), "k_slice_sum launch");
}

int launch_slice_sum(modp_ctx *ctx, const GemmArgs &g, int epi, const int32_t *P, int S)
{
    if (g.kc == EK_F64) return launch_slice_sum_t<double>(ctx, g, epi, P, S);
    if (g.kc == EK_F32) return launch_slice_sum_t<float>(ctx, g, epi, P, S);
    return launch_slice_sum_t<int32_t>(ctx, g, epi, P, S);
}

// slice results go to a workspace of S * m * n int32 residues
int32_t *slice_buffer(modp_ctx *ctx, const GemmArgs &g, int S, int *st)
{
    return (int32_t *)ws_get(ctx, WS_SLICES, sizeof(int32_t) * (size_t)S * (size_t)g.m * (size_t)g.n, st);
}

template <typename TA>
int split_a_t(modp_ctx *ctx, const GemmArgs &g, const void *A, int64_t kbeg, int64_t kc, int8_t *Al,
              int64_t Mp, int64_t Kp, int L, int32_t top)
{
    dim3 grid((unsigned)(Kp / 64), (unsigned)(Mp / 64));
    const TA *Ap = (const TA *)A + kbeg * g.lda;
    k_split_a<TA><<<grid, 256, 0, ctx->stream>>>(g.m, kc, Ap, g.lda, Al, Mp, Kp, L, top, g.mp, g.skip);
    return hip_check(ctx, hipGetLastError(), "k_split_a launch");
}

template <typename TB>
int split_b_t(modp_ctx *ctx, const GemmArgs &g, const void *B, int64_t kbeg, int64_t kc, int8_t *Bl,
              int64_t Np, int64_t Kp, int L, int32_t top)
{
    int64_t total = Np * (Kp / 4);
    unsigned grid = (unsigned)((total + 255) / 256);
    const TB *Bp = (const TB *)B + kbeg;
    k_split_b<TB><<<grid, 256, 0, ctx->stream>>>(kc, g.n, Bp, g.ldb, Bl, Np, Kp, L, top, g.mp, g.skip);
    return hip_check(ctx, hipGetLastError(), "k_split_b launch");
}

int launch_limb(modp_ctx *ctx, const GemmArgs &g, int L, int epi, const int8_t *Al, int64_t Mp,
                const int8_t *Bl, int64_t Np, int64_t Kp, const LimbW &lw, int tm, int tn, const Slices &sl)
{
    switch (L) {
    case 1: return launch_limb_impl<1>(ctx, g, epi, Al, Mp, Bl, Np, Kp, lw, tm, tn, sl);
    case 2: return launch_limb_impl<2>(ctx, g, epi, Al, Mp, Bl, Np, Kp, lw, tm, tn, sl);
    case 3: return launch_limb_impl<3>(ctx, g, epi, Al, Mp, Bl, Np, Kp, lw, tm, tn, sl);
    default: return launch_limb_impl<4>(ctx, g, epi, Al, Mp, Bl, Np, Kp, lw, tm, tn, sl);
    }
}

int64_t tiles_of(const GemmArgs &g, int tm, int tn) { return ((g.m + tm - 1) / tm) * ((g.n + tn - 1) / tn); }

// slices of a K extent: S0 requested, each a multiple of `align` and at
// least `minlen` long
Slices make_slices(int64_t kext, int S0, int64_t align, int64_t minlen)
{
    Slices sl;
    if (S0 <= 1 || kext <= minlen) {
        sl.S = 1;
        sl.kslice = kext > 0 ? kext : 1;
        return sl;
    }
    int64_t ks = round_up((kext + S0 - 1) / S0, align);
    if (ks < minlen) ks = round_up(minlen, align);
    sl.S = (int)((kext + ks - 1) / ks);
    sl.kslice = ks;
    return sl;
}

int gemm_plain_cfg(modp_ctx *ctx, const GemmArgs &g, const GemmConfig &cfg)
{
    if (g.m <= 0 || g.n <= 0) return MODP_OK;
    const int R = f64_reduce_interval(g.mp.p);
    if (R < 1) return set_error(ctx, MODP_ERR_MODULUS_RANGE, "modulus %lld too large for the FP64 form", (long long)g.mp.p);
    const int tile = cfg.tile_m == 128 ? 128 : 64;
    Slices sl = make_slices(g.k, cfg.ksplit, PBK, 64);
    if (sl.S == 1) {
        sl.kslice = g.k > 0 ? g.k : 1;
        return launch_f64(ctx, g, R, tile, sl);
    }
    int st = MODP_OK;
    int32_t *P = slice_buffer(ctx, g, sl.S, &st);
    if (!P) return st;
    GemmArgs gs = g;
    gs.C = P; gs.ldc = g.m; gs.kc = EK_I32; gs.epi = EPI_STORE;
    sl.cslice = g.m * g.n;
    MODP_TRY(launch_f64(ctx, gs, R, tile, sl));
    return launch_slice_sum(ctx, g, g.epi, P, sl.S);
}

int gemm_limb_cfg(modp_ctx *ctx, const GemmArgs &g, const GemmConfig &cfg)
{
    if (g.m <= 0 || g.n <= 0) return MODP_OK;
    const int L = limb_count(g.mp.p);
    if (L > 4) return set_error(ctx, MODP_ERR_MODULUS_RANGE, "modulus %lld needs more than 4 limbs", (long long)g.mp.p);
    // largest k-chunk whose shift sums fit int32: a digit product is at most
    // 128^2 in magnitude and shift L-1 collects L of them per k, so
    // L * K * 2^14 <= 2^31 - 1
    const int64_t pairs = L;
    int64_t kmax = ((int64_t)2147483647 / (pairs * 16384)) / LBK * LBK;
    // residues above top are represented as r - p (see digit_rep)
    int64_t S = 0;
    for (int u = 0; u < L; ++u) S = S * 256 + 1;
    const int32_t top = (int32_t)(127 * S > 2147483647 ? 2147483647 : 127 * S);
    const int tm = cfg.tile_m == 128 ? 128 : 64;
    const int tn = (cfg.tile_m == 128 && cfg.tile_n == 128 && L <= 2) ? 128 : 64;
    const int64_t Mp = round_up(g.m, tm), Np = round_up(g.n, tn);
    LimbW lw;
    {
        double w = 1.0;
        for (int u = 0; u < 7; ++u) {
            lw.w[u] = w;
            w = fmod(w * 256.0, (double)g.mp.p);
        }
    }
    if (g.k <= 0 && g.epi != EPI_STORE) return MODP_OK;  // C unchanged
    int64_t kdone = 0;
    bool first = true;
    do {
        int64_t kc = g.k - kdone;
        if (kc > kmax) kc = kmax;
        const int64_t Kp = round_up(kc > 0 ? kc : 1, LBK);
        int st = MODP_OK;
        int8_t *Al = (int8_t *)ws_get(ctx, WS_ALIMB, (size_t)L * Mp * Kp, &st);
        if (!Al) return st;
        int8_t *Bl = (int8_t *)ws_get(ctx, WS_BLIMB, (size_t)L * Np * Kp, &st);
        if (!Bl) return st;
        if (g.ka == EK_F64) MODP_TRY(split_a_t<double>(ctx, g, g.A, kdone, kc, Al, Mp, Kp, L, top));
        else if (g.ka == EK_F32) MODP_TRY(split_a_t<float>(ctx, g, g.A, kdone, kc, Al, Mp, Kp, L, top));
        else MODP_TRY(split_a_t<int32_t>(ctx, g, g.A, kdone, kc, Al, Mp, Kp, L, top));
        if (g.kb == EK_F64) MODP_TRY(split_b_t<double>(ctx, g, g.B, kdone, kc, Bl, Np, Kp, L, top));
        else if (g.kb == EK_F32) MODP_TRY(split_b_t<float>(ctx, g, g.B, kdone, kc, Bl, Np, Kp, L, top));
        else MODP_TRY(split_b_t<int32_t>(ctx, g, g.B, kdone, kc, Bl, Np, Kp, L, top));
        int epi = g.epi;
        if (!first && epi == EPI_STORE) epi = EPI_ADD;
        Slices sl = make_slices(Kp, cfg.ksplit, LBK, 256);
        if (sl.S == 1) {
            sl.kslice = Kp;
            MODP_TRY(launch_limb(ctx, g, L, epi, Al, Mp, Bl, Np, Kp, lw, tm, tn, sl));
        } else {
            int32_t *P = slice_buffer(ctx, g, sl.S, &st);
            if (!P) return st;
            GemmArgs gs = g;
            gs.C = P; gs.ldc = g.m; gs.kc = EK_I32;
            sl.cslice = g.m * g.n;
            MODP_TRY(launch_limb(ctx, gs, L, EPI_STORE, Al, Mp, Bl, Np, Kp, lw, tm, tn, sl));
            MODP_TRY(launch_slice_sum(ctx, g, epi, P, sl.S));
        }
        kdone += kc;
        first = false;
    } while (kdone < g.k);
    return MODP_OK;
}

}  // namespace

int limb_count(int64_t p)
{
    // signed base-256 digits: L digits represent every residue class iff p <= 256^L
    int L = 1;
    while (L < 8 && p > ((int64_t)1 << (8 * L))) ++L;
    return L;
}

int choose_form(int64_t m, int64_t n, int64_t k, int64_t p)
{
    (void)m; (void)n; (void)p;
    // the limb form pays a split pass over A and B: worth it once the
    // contraction is deep enough for MFMA time to dominate
    if (k >= 256) return MODP_FORM_LIMB_I8;
    return MODP_FORM_PLAIN_F64;
}

bool gemm_config_valid(const GemmArgs &g, const GemmConfig &cfg)
{
    if (cfg.ksplit < 1 || cfg.ksplit > 16) return false;
    const bool sq64 = cfg.tile_m == 64 && cfg.tile_n == 64;
    const bool sq128 = cfg.tile_m == 128 && cfg.tile_n == 128;
    const bool r128x64 = cfg.tile_m == 128 && cfg.tile_n == 64;
    if (cfg.form == MODP_FORM_PLAIN_F64) return (sq64 || sq128) && f64_reduce_interval(g.mp.p) >= 1;
    // 128x128 limb tiles hold 2L-1 accumulators of a 64x64 block per wave:
    // register room for L <= 2 only
    if (cfg.form == MODP_FORM_LIMB_I8) {
        const int L = limb_count(g.mp.p);
        return (sq64 || r128x64 || (sq128 && L <= 2)) && L <= 4;
    }
    return false;
}

GemmConfig gemm_default_config(const modp_ctx *ctx, const GemmArgs &g, int form)
{
    GemmConfig c;
    c.form = form == MODP_FORM_AUTO ? choose_form(g.m, g.n, g.k, g.mp.p) : form;
    c.tile_m = c.tile_n = 64;
    // larger tiles while the grid still fills the chip: measured best at
    // n = 8192 are 128x128 (plain; limb with one digit) and 128x64 (limb with
    // two digits); three and four digits gain nothing over 64x64
    if (tiles_of(g, 128, 128) >= 2 * ctx->num_cus) {
        const int L = limb_count(g.mp.p);
        if (c.form == MODP_FORM_PLAIN_F64 || L == 1) {
            c.tile_m = c.tile_n = 128;
        } else if (L == 2) {
            c.tile_m = 128;
        }
    }
    // a grid far below one workgroup per CU: cut K into concurrent slices
    const int64_t t = tiles_of(g, c.tile_m, c.tile_n);
    c.ksplit = 1;
    if (2 * t <= ctx->num_cus && g.k >= 2048) {
        int64_t S = ctx->num_cus / (t > 0 ? t : 1);
        if (S > 8) S = 8;
        while (S > 1 && g.k / S < 512) --S;
        c.ksplit = (int)S;
    }
    return c;
}

int gemm_run(modp_ctx *ctx, const GemmArgs &g, const GemmConfig &cfg)
{
    if (cfg.form == MODP_FORM_LIMB_I8) return gemm_limb_cfg(ctx, g, cfg);
    return gemm_plain_cfg(ctx, g, cfg);
}

int gemm_plain_f64(modp_ctx *ctx, const GemmArgs &g)
{
    return gemm_plain_cfg(ctx, g, gemm_default_config(ctx, g, MODP_FORM_PLAIN_F64));
}

int gemm_limb_i8(modp_ctx *ctx, const GemmArgs &g)
{
    return gemm_limb_cfg(ctx, g, gemm_default_config(ctx, g, MODP_FORM_LIMB_I8));
}
