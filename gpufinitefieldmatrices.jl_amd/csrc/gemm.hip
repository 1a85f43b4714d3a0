// gemm.hip -- mod-p GEMM cores for gfx950.
//
// Plain form: FP64 MFMA (v_mfma_f64_16x16x4f64) on balanced residues of
// magnitude at most (p+1)/2, with the accumulators reduced every R k4-steps
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
template <typename TA, typename TB, typename TC, int BM, int BN, int EPI, int RED>
__global__ __launch_bounds__(256, 2) void k_gemm_f64(int64_t m, int64_t n, int64_t k,
                                                  const TA *__restrict__ A, int64_t lda,
                                                  const TB *__restrict__ B, int64_t ldb,
                                                  TC *__restrict__ C, int64_t ldc, ModP mp, int rq,
                                                  int64_t kslice, int64_t cslice,
                                                  const int *__restrict__ skip)
{
    if (skip && *skip) return;
    int64_t tm, tn;
    if (!tile_of_block((m + BM - 1) / BM, (n + BN - 1) / BN, &tm, &tn)) return;
    const int64_t kb = (int64_t)blockIdx.y * kslice;
    const int64_t klen = k - kb < kslice ? k - kb : kslice;
    __shared__ F64Smem<BM, BN> sm;
    gemm_f64_tile<TA, TB, TC, BM, BN, EPI, RED>(m, n, klen, A + kb * lda, lda, B + kb, ldb, C + blockIdx.y * cslice,
                                                ldc, mp, rq, tm * BM, tn * BN, sm);
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

// y = A x mod p (A m x k column-major, x and y contiguous): one thread per
// row, blockIdx.y a K slice.  Balanced residues; the FP64 accumulator is
// reduced every R1 products, |a x| <= floor(p/2)^2, so it stays exact.
// A slice writes its residues to P + y*m when S > 1, else straight to y.
template <typename T>
__global__ __launch_bounds__(256) void k_gemv(int64_t m, int64_t k, const T *__restrict__ A, int64_t lda,
                                              const T *__restrict__ x, T *__restrict__ y, int32_t *__restrict__ P,
                                              int64_t kslice, int R1, ModP mp)
{
    __shared__ double xs[256];
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t kb = (int64_t)blockIdx.y * kslice;
    const int64_t ke = k - kb < kslice ? k : kb + kslice;
    const double hp = (double)mp.half;
    double acc = 0.0;
    int cnt = 0;
    for (int64_t k0 = kb; k0 < ke; k0 += 256) {
        const int64_t kk = k0 + threadIdx.x;
        double xv = 0.0;
        if (kk < ke) {
            xv = (double)to_res<T>(x[kk], mp);
            xv = xv > hp ? xv - mp.pd : xv;
        }
        __syncthreads();
        xs[threadIdx.x] = xv;
        __syncthreads();
        const int kn = ke - k0 < 256 ? (int)(ke - k0) : 256;
        if (r < m) {
            const T *a = A + r + k0 * lda;
#pragma unroll 8
            for (int q = 0; q < kn; ++q) {
                double av = (double)to_res<T>(a[q * lda], mp);
                av = av > hp ? av - mp.pd : av;
                acc = fma(av, xs[q], acc);
                if (++cnt == R1) {
                    cnt = 0;
                    acc = fma(-rint(acc * mp.pinv), mp.pd, acc);
                }
            }
        }
    }
    if (r < m) {
        const uint32_t v = (uint32_t)dmod(acc, mp);
        if (P) P[(int64_t)blockIdx.y * m + r] = (int32_t)v;
        else y[r] = from_res<T>(v);
    }
}

int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// ---- dispatch helpers -----------------------------------------------------

template <typename TA, typename TB, typename TC, int BM, int BN, int EPI>
int launch_f64_e(modp_ctx *ctx, const GemmArgs &g, int R, const Slices &sl, const dim3 &grid)
{
    const TA *A = (const TA *)g.A;
    const TB *B = (const TB *)g.B;
    TC *C = (TC *)g.C;
    int rq = 1;
    const int red = f64_reduce_mode(R, sl.kslice < g.k ? sl.kslice : g.k, &rq);
#define F64_LAUNCH(RED_)                                                                                      \
    k_gemm_f64<TA, TB, TC, BM, BN, EPI, RED_><<<grid, 256, 0, ctx->stream>>>(g.m, g.n, g.k, A, g.lda, B, g.ldb, C, \
                                                                            g.ldc, g.mp, rq, sl.kslice,          \
                                                                            sl.cslice, g.skip)
    switch (red) {
    case RED_NONE: F64_LAUNCH(RED_NONE); break;
    case RED_STAGE: F64_LAUNCH(RED_STAGE); break;
    case RED_2: F64_LAUNCH(RED_2); break;
    default: F64_LAUNCH(RED_STEP); break;
    }
#undef F64_LAUNCH
    return hip_check(ctx, hipGetLastError(), "k_gemm_f64 launch");
}

template <typename TA, typename TC, int EPI>
int launch_f64_tile(modp_ctx *ctx, const GemmArgs &g, int R, int tile, const Slices &sl)
{
    TimedLaunch tl(ctx, "gemm_f64");
    if (tile == 128) {
        const dim3 grid(xcd_grid(((g.n + 127) / 128) * ((g.m + 127) / 128)), (unsigned)sl.S);
        return launch_f64_e<TA, TA, TC, 128, 128, EPI>(ctx, g, R, sl, grid);
    }
    const dim3 grid(xcd_grid(((g.n + 63) / 64) * ((g.m + 63) / 64)), (unsigned)sl.S);
    return launch_f64_e<TA, TA, TC, 64, 64, EPI>(ctx, g, R, sl, grid);
}

// The operand / result / epilogue combinations the library uses: user
// products (C of the operands' type, or int32 K-slice partials), and the
// int32 working arrays of elimination and inversion (subtracting updates,
// products into the caller's storage type).
int launch_f64(modp_ctx *ctx, const GemmArgs &g, int R, int tile, const Slices &sl)
{
    if (g.ka == g.kb) {
        if (g.epi == EPI_STORE) {
            if (g.ka == EK_F64 && g.kc == EK_F64) return launch_f64_tile<double, double, EPI_STORE>(ctx, g, R, tile, sl);
            if (g.ka == EK_F64 && g.kc == EK_I32) return launch_f64_tile<double, int32_t, EPI_STORE>(ctx, g, R, tile, sl);
            if (g.ka == EK_F32 && g.kc == EK_F32) return launch_f64_tile<float, float, EPI_STORE>(ctx, g, R, tile, sl);
            if (g.ka == EK_F32 && g.kc == EK_I32) return launch_f64_tile<float, int32_t, EPI_STORE>(ctx, g, R, tile, sl);
            if (g.ka == EK_I32 && g.kc == EK_I32) return launch_f64_tile<int32_t, int32_t, EPI_STORE>(ctx, g, R, tile, sl);
            if (g.ka == EK_I32 && g.kc == EK_F64) return launch_f64_tile<int32_t, double, EPI_STORE>(ctx, g, R, tile, sl);
            if (g.ka == EK_I32 && g.kc == EK_F32) return launch_f64_tile<int32_t, float, EPI_STORE>(ctx, g, R, tile, sl);
        } else if (g.epi == EPI_SUB && g.ka == EK_I32 && g.kc == EK_I32) {
            return launch_f64_tile<int32_t, int32_t, EPI_SUB>(ctx, g, R, tile, sl);
        }
    }
    return set_error(ctx, MODP_ERR_DTYPE, "FP64 GEMM: unsupported operand combination (%d, %d, %d, epilogue %d)",
                     g.ka, g.kb, g.kc, g.epi);
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

// ext: caller-provided workspace for the digit planes (no allocation); the
// planes of each K chunk must fit in ext->bytes and the launch has one K slice
int gemm_limb_cfg(modp_ctx *ctx, const GemmArgs &g, const GemmConfig &cfg, const LimbScratch *ext = nullptr)
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
        int8_t *Al, *Bl;
        if (ext) {
            const size_t need = (size_t)L * (size_t)(Mp + Np) * (size_t)Kp;
            if (need > ext->bytes || cfg.ksplit != 1)
                return set_error(ctx, MODP_ERR_ARG, "digit planes of %zu bytes exceed the %zu-byte workspace", need,
                                 ext->bytes);
            Al = ext->base;
            Bl = ext->base + (size_t)L * Mp * Kp;
        } else {
            Al = (int8_t *)ws_get(ctx, WS_ALIMB, (size_t)L * Mp * Kp, &st);
            if (!Al) return st;
            Bl = (int8_t *)ws_get(ctx, WS_BLIMB, (size_t)L * Np * Kp, &st);
            if (!Bl) return st;
        }
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
    if (cfg.form == MODP_FORM_RNS_I8) {
        const bool sq256 = cfg.tile_m == 256 && cfg.tile_n == 256;
        return (sq128 || sq256) && cfg.ksplit == 1 && rns_planes(g.mp.p, g.k < RNS_KCHUNK ? g.k : RNS_KCHUNK) > 0;
    }
    // 128x128 limb tiles hold 2L-1 accumulators of a 64x64 block per wave:
    // register room for L <= 2 only
    if (cfg.form == MODP_FORM_LIMB_I8) {
        const int L = limb_count(g.mp.p);
        return (sq64 || r128x64 || (sq128 && L <= 2)) && L <= 4;
    }
    return false;
}

// the RNS form's r moduli against the limb form's L^2 digit products: it
// pays when it issues at most three quarters of the limb form's MFMA work,
// and with one plane (p <= 255) against one digit product, whose core runs
// at a lower fraction of the MFMA rate
static bool rns_pays(const GemmArgs &g)
{
    const int L = limb_count(g.mp.p);
    const int r = rns_planes(g.mp.p, g.k < RNS_KCHUNK ? g.k : RNS_KCHUNK);
    return r > 0 && (r == 1 || 4 * r <= 3 * L * L);
}

GemmConfig gemm_default_config(const modp_ctx *ctx, const GemmArgs &g, int form)
{
    GemmConfig c;
    c.form = form == MODP_FORM_AUTO ? choose_form(g.m, g.n, g.k, g.mp.p) : form;
    // (one plane only with at least two 256 x 256 output tiles per CU:
    // measured at p = 127, the one-digit form is faster at 4096^2 -- 646
    // against 592 TOPS -- and the one plane at 8192^2, 1060 against 967)
    if (form == MODP_FORM_AUTO && c.form == MODP_FORM_LIMB_I8 && rns_pays(g) &&
        (rns_planes(g.mp.p, g.k < RNS_KCHUNK ? g.k : RNS_KCHUNK) > 1 || tiles_of(g, 256, 256) >= 2 * ctx->num_cus))
        c.form = MODP_FORM_RNS_I8;
    c.ksplit = 1;
    if (c.form == MODP_FORM_RNS_I8) {
        // 256x256 tiles while the grid (tiles x moduli) still holds two
        // workgroups per CU
        const int r = rns_planes(g.mp.p, g.k < RNS_KCHUNK ? g.k : RNS_KCHUNK);
        c.tile_m = c.tile_n = tiles_of(g, 256, 256) * r >= 2 * ctx->num_cus ? 256 : 128;
        return c;
    }
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
    // a grid of at most one workgroup per CU: cut K into concurrent slices,
    // up to about four workgroups per CU, each slice at least 256 deep
    // (1024^2 plain: 0.238 ms unsplit, 0.142 / 0.100 / 0.106 ms with 2 / 4 /
    // 8 slices of 64x64 tiles)
    const int64_t t = tiles_of(g, c.tile_m, c.tile_n);
    c.ksplit = 1;
    if (t <= ctx->num_cus && g.k >= 1024) {
        int64_t S = 4 * ctx->num_cus / (t > 0 ? t : 1);
        if (S > 8) S = 8;
        while (S > 1 && g.k / S < 256) --S;
        c.ksplit = (int)S;
    }
    return c;
}

int gemm_run(modp_ctx *ctx, const GemmArgs &g, const GemmConfig &cfg)
{
    if (cfg.form == MODP_FORM_LIMB_I8) return gemm_limb_cfg(ctx, g, cfg);
    if (cfg.form == MODP_FORM_RNS_I8) return gemm_rns_run(ctx, g, cfg.tile_m);
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

int gemm_rns_i8(modp_ctx *ctx, const GemmArgs &g)
{
    return gemm_run(ctx, g, gemm_default_config(ctx, g, MODP_FORM_RNS_I8));
}

size_t gemm_limb_scratch_bytes(const modp_ctx *ctx, const GemmArgs &g)
{
    GemmConfig cfg = gemm_default_config(ctx, g, MODP_FORM_LIMB_I8);
    const int L = limb_count(g.mp.p);
    const int tm = cfg.tile_m == 128 ? 128 : 64;
    const int tn = (cfg.tile_m == 128 && cfg.tile_n == 128 && L <= 2) ? 128 : 64;
    const int64_t kmax = ((int64_t)2147483647 / ((int64_t)L * 16384)) / LBK * LBK;
    const int64_t kc = g.k < kmax ? g.k : kmax;
    const int64_t Kp = round_up(kc > 0 ? kc : 1, LBK);
    return (size_t)L * (size_t)(round_up(g.m, tm) + round_up(g.n, tn)) * (size_t)Kp;
}

int gemm_limb_i8_in(modp_ctx *ctx, const GemmArgs &g, const LimbScratch &ws)
{
    GemmConfig cfg = gemm_default_config(ctx, g, MODP_FORM_LIMB_I8);
    cfg.ksplit = 1;
    return gemm_limb_cfg(ctx, g, cfg, &ws);
}

int gemv_run(modp_ctx *ctx, int kind, int64_t m, int64_t k, const void *A, int64_t lda, const void *x, void *y,
             int64_t p)
{
    if (m <= 0) return MODP_OK;
    const ModP mp = make_modp(p);
    const double h = (double)(p / 2);
    const double r1 = h > 0 ? floor((9007199254740992.0 - (double)p) / (h * h)) : 1e9;
    const int R1 = r1 > 1e6 ? 1000000 : (int)r1;
    if (R1 < 1) return set_error(ctx, MODP_ERR_MODULUS_RANGE, "modulus %lld too large for the FP64 form", (long long)p);
    const int64_t rb = (m + 255) / 256;
    // K slices until the grid holds about two workgroups per CU
    int64_t S = (2 * (int64_t)ctx->num_cus + rb - 1) / rb;
    if (S > (k + 511) / 512) S = (k + 511) / 512;
    if (S < 1) S = 1;
    const int64_t kslice = round_up((k + S - 1) / S > 0 ? (k + S - 1) / S : 1, 256);
    S = k > 0 ? (k + kslice - 1) / kslice : 1;
    int st = MODP_OK;
    int32_t *P = nullptr;
    if (S > 1) {
        P = (int32_t *)ws_get(ctx, WS_SLICES, sizeof(int32_t) * (size_t)S * (size_t)m, &st);
        if (!P) return st;
    }
    const dim3 grid((unsigned)rb, (unsigned)S);
    {
        TimedLaunch tl(ctx, "gemv");
        if (kind == EK_F64)
            k_gemv<double><<<grid, 256, 0, ctx->stream>>>(m, k, (const double *)A, lda, (const double *)x, (double *)y,
                                                          P, kslice, R1, mp);
        else
            k_gemv<float><<<grid, 256, 0, ctx->stream>>>(m, k, (const float *)A, lda, (const float *)x, (float *)y, P,
                                                         kslice, R1, mp);
        HIP_TRY(ctx, hipGetLastError());
    }
    if (S == 1) return MODP_OK;
    GemmArgs g;
    g.m = m; g.n = 1; g.k = k;
    g.C = y; g.ldc = m; g.kc = kind;
    g.epi = EPI_STORE;
    g.mp = mp;
    return launch_slice_sum(ctx, g, EPI_STORE, P, (int)S);
}
