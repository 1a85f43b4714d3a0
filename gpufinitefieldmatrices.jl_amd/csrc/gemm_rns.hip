// gemm_rns.hip -- host side of the multi-modular (RNS) int8 form: the moduli
// for (p, k), the K chunks, the workspace and the three launches (split,
// one int8 GEMM per modulus, combine).  Kernels in gemm_rns.h.
#include "gemm_rns.h"

namespace {

using rns::Moduli;
using rns::NMOD_MAX;
using rns::QLIST;
typedef unsigned __int128 u128;

int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

int64_t inv_mod_small(int64_t a, int64_t q)
{
    // extended Euclid (Knuth TAOCP vol. 2 4.5.2), 0 < a < q, gcd(a, q) = 1
    int64_t t = 0, nt = 1, r = q, nr = a % q;
    while (nr != 0) {
        const int64_t qq = r / nr;
        int64_t x = t - qq * nt; t = nt; nt = x;
        x = r - qq * nr; r = nr; nr = x;
    }
    return t < 0 ? t + q : t;
}

}  // namespace

// Smallest r with prod_{j<r} q_j > 4 k floor(p/2)^2 (so |C_int| / M < 1/4),
// or 0 when twelve moduli are not enough.
int rns_count(int64_t p, int64_t k)
{
    const u128 h = (u128)(p / 2);
    const u128 bound = (u128)4 * (u128)(k > 0 ? k : 1) * h * h;
    u128 M = 1;
    for (int j = 0; j < NMOD_MAX; ++j) {
        M *= (u128)QLIST[j];
        if (M > bound) return j + 1;
    }
    return 0;
}

// 2 <= p <= 255: one int8 product with the modulus p itself.  Balanced
// residues (|x| <= 127) keep a K chunk's int32 sums exact, and the plane
// holds C mod p, so no Chinese remaindering is needed.
bool rns_single(int64_t p) { return p >= 2 && p <= 255; }

int rns_planes(int64_t p, int64_t k) { return rns_single(p) ? 1 : rns_count(p, k); }

// the single plane as a one-modulus system with M = q_0 = p: y_0 = 1,
// w_0 = 1 and M mod p = 0, so the recombination returns the plane's residue
static Moduli single_modulus(int64_t p)
{
    Moduli md{};
    md.r = 1;
    md.single = 1;
    md.q[0] = (int)p;
    md.y[0] = 1;
    md.w[0] = 1.0;
    md.invq[0] = 1.0 / (double)p;
    md.wm = 0.0;
    return md;
}

static Moduli make_moduli(int64_t p, int r)
{
    Moduli md{};
    md.r = r;
    u128 M = 1;
    for (int j = 0; j < r; ++j) M *= (u128)QLIST[j];
    for (int j = 0; j < r; ++j) {
        const int64_t q = QLIST[j];
        const u128 Mj = M / (u128)q;
        md.q[j] = (int)q;
        md.y[j] = (int)inv_mod_small((int64_t)(Mj % (u128)q), q);
        md.w[j] = (double)(int64_t)(Mj % (u128)p);
        md.invq[j] = 1.0 / (double)q;
    }
    md.wm = (double)(int64_t)(M % (u128)p);
    return md;
}

namespace {

template <class SH>
int launch_core(modp_ctx *ctx, const int8_t *X, const int8_t *Y, int64_t Np, int64_t Mp, int64_t Kp, uint8_t *E,
                const Moduli &md)
{
    const dim3 grid(xcd_grid((Np / SH::TN) * (Mp / SH::TM)), (unsigned)md.r);
    TimedLaunch tl(ctx, "gemm_rns");
    rns::k_rns_gemm2<SH><<<grid, SH::NT, 0, ctx->stream>>>(X, Y, Np, Mp, Kp, E, md);
    return hip_check(ctx, hipGetLastError(), "k_rns_gemm launch");
}

template <typename T>
int split_rows_t(modp_ctx *ctx, const GemmArgs &g, int64_t kbeg, int64_t kc, int8_t *Yp, int64_t Mp, int64_t Kp,
                 const Moduli &md)
{
    const dim3 grid((unsigned)(Kp / 128), (unsigned)(Mp / 64));
    const T *A = (const T *)g.A + kbeg * g.lda;
    rns::k_split_rows<T><<<grid, 256, 0, ctx->stream>>>(g.m, kc, A, g.lda, Yp, Mp, Kp, g.mp, md, g.skip);
    return hip_check(ctx, hipGetLastError(), "k_split_rows launch");
}

template <typename T>
int split_cols_t(modp_ctx *ctx, const GemmArgs &g, int64_t kbeg, int64_t kc, int8_t *Xp, int64_t Np, int64_t Kp,
                 const Moduli &md)
{
    const int64_t total = Np * (Kp / 16);
    const T *B = (const T *)g.B + kbeg;
    rns::k_split_cols<T><<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(kc, g.n, B, g.ldb, Xp, Np, Kp,
                                                                                   g.mp, md, g.skip);
    return hip_check(ctx, hipGetLastError(), "k_split_cols launch");
}

template <typename TC>
int combine_t(modp_ctx *ctx, const GemmArgs &g, int epi, const uint8_t *E, int64_t Mp, int64_t plane,
              const Moduli &md)
{
    TC *C = (TC *)g.C;
    if (md.single) {
        // one wave per 1024 rows of a column, four per workgroup
        const unsigned g1 = (unsigned)((((g.m + 1023) / 1024) * g.n + 3) / 4);
        if (epi == EPI_STORE)
            rns::k_combine1<TC, EPI_STORE><<<g1, 256, 0, ctx->stream>>>(g.m, g.n, E, Mp, C, g.ldc, g.mp, g.skip);
        else if (epi == EPI_SUB)
            rns::k_combine1<TC, EPI_SUB><<<g1, 256, 0, ctx->stream>>>(g.m, g.n, E, Mp, C, g.ldc, g.mp, g.skip);
        else
            rns::k_combine1<TC, EPI_ADD><<<g1, 256, 0, ctx->stream>>>(g.m, g.n, E, Mp, C, g.ldc, g.mp, g.skip);
        return hip_check(ctx, hipGetLastError(), "k_combine1 launch");
    }
    const unsigned grid = (unsigned)((((g.m + 3) / 4) * g.n + 255) / 256);
    // 16-byte stores of four rows need every column start 16-byte aligned
    const bool vec = sizeof(TC) == 8 && (g.ldc & 1) == 0 && ((uintptr_t)C & 15) == 0;
    if (epi == EPI_STORE)
        rns::k_combine<TC, EPI_STORE><<<grid, 256, 0, ctx->stream>>>(g.m, g.n, E, Mp, plane, md, C, g.ldc, g.mp, g.skip,
                                                                      vec);
    else if (epi == EPI_SUB)
        rns::k_combine<TC, EPI_SUB><<<grid, 256, 0, ctx->stream>>>(g.m, g.n, E, Mp, plane, md, C, g.ldc, g.mp, g.skip,
                                                                    vec);
    else
        rns::k_combine<TC, EPI_ADD><<<grid, 256, 0, ctx->stream>>>(g.m, g.n, E, Mp, plane, md, C, g.ldc, g.mp, g.skip,
                                                                    vec);
    return hip_check(ctx, hipGetLastError(), "k_combine launch");
}

// output tile (n side x m side) -> workgroup shape: 256x256 with 2x4 waves of
// 128x64 (two waves per SIMD; measured 4.54 ms at 8192^3 with nine moduli
// against 5.00 ms for 2x2 waves of 128x128 with the accumulators in AGPRs)
// or 128x128 with 2x2 waves of 64x64
int core_by_tile(modp_ctx *ctx, int tile, const int8_t *X, const int8_t *Y, int64_t Np, int64_t Mp, int64_t Kp,
                 uint8_t *E, const Moduli &md)
{
    if (tile == 256) return launch_core<rns::Shape<256, 256, 2, 4, 128>>(ctx, X, Y, Np, Mp, Kp, E, md);
    return launch_core<rns::Shape<128, 128, 2, 2, 128>>(ctx, X, Y, Np, Mp, Kp, E, md);
}

}  // namespace

int gemm_rns_run(modp_ctx *ctx, const GemmArgs &g, int tile)
{
    if (g.m <= 0 || g.n <= 0) return MODP_OK;
    if (g.k <= 0 && g.epi != EPI_STORE) return MODP_OK;
    // K chunks of at most RNS_KCHUNK keep the int32 sums exact
    // (|residue products| <= 2^14 and RNS_KCHUNK * 2^14 < 2^31)
    const int64_t kc_max = RNS_KCHUNK;
    const int64_t T = tile == 256 ? 256 : 128;
    const int64_t Mp = round_up(g.m, T), Np = round_up(g.n, T);
    int64_t kdone = 0;
    bool first = true;
    do {
        int64_t kc = g.k - kdone;
        if (kc > kc_max) kc = kc_max;
        const int r = rns_planes(g.mp.p, kc > 0 ? kc : 1);
        if (r == 0) return set_error(ctx, MODP_ERR_MODULUS_RANGE, "modulus %lld too large for the RNS form",
                                     (long long)g.mp.p);
        const Moduli md = rns_single(g.mp.p) ? single_modulus(g.mp.p) : make_moduli(g.mp.p, r);
        const int64_t Kp = round_up(kc > 0 ? kc : 1, 128);
        int st = MODP_OK;
        int8_t *Yp = (int8_t *)ws_get(ctx, WS_ALIMB, (size_t)r * Mp * Kp, &st);
        if (!Yp) return st;
        int8_t *Xp = (int8_t *)ws_get(ctx, WS_BLIMB, (size_t)r * Np * Kp, &st);
        if (!Xp) return st;
        uint8_t *E = (uint8_t *)ws_get(ctx, WS_SLICES, (size_t)r * Np * Mp, &st);
        if (!E) return st;
        if (g.ka == EK_F64) MODP_TRY(split_rows_t<double>(ctx, g, kdone, kc, Yp, Mp, Kp, md));
        else if (g.ka == EK_F32) MODP_TRY(split_rows_t<float>(ctx, g, kdone, kc, Yp, Mp, Kp, md));
        else MODP_TRY(split_rows_t<int32_t>(ctx, g, kdone, kc, Yp, Mp, Kp, md));
        if (g.kb == EK_F64) MODP_TRY(split_cols_t<double>(ctx, g, kdone, kc, Xp, Np, Kp, md));
        else if (g.kb == EK_F32) MODP_TRY(split_cols_t<float>(ctx, g, kdone, kc, Xp, Np, Kp, md));
        else MODP_TRY(split_cols_t<int32_t>(ctx, g, kdone, kc, Xp, Np, Kp, md));
        MODP_TRY(core_by_tile(ctx, (int)T, Xp, Yp, Np, Mp, Kp, E, md));
        int epi = g.epi;
        if (!first && epi == EPI_STORE) epi = EPI_ADD;
        if (g.kc == EK_F64) MODP_TRY(combine_t<double>(ctx, g, epi, E, Mp, Np * Mp, md));
        else if (g.kc == EK_F32) MODP_TRY(combine_t<float>(ctx, g, epi, E, Mp, Np * Mp, md));
        else MODP_TRY(combine_t<int32_t>(ctx, g, epi, E, Mp, Np * Mp, md));
        kdone += kc;
        first = false;
    } while (kdone < g.k);
    return MODP_OK;
}
