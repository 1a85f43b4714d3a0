// gemm_limb.h -- limb form of the mod-p GEMM (gfx950 int8 MFMA).
//
// Each residue r in [0,p) is written in balanced base 256,
// r == sum_u d_u 256^u (mod p), with signed digits d_u in [-128,127] held in
// int8; L = ceil(log256 p) digits (p <= 256^L) cover every residue class.
// The L*L digit products run on v_mfma_i32_16x16x64_i8 with int32
// accumulation, products of equal shift s+t sharing one accumulator; K is cut
// into chunks so that no int32 partial sum can overflow; the 2L-1 shift sums
// are reduced mod p and recombined with weights 256^u mod p in FP64.
//
// The kernel templates live here; gemm_limb_l{1,2,3,4}.hip instantiate the
// launchers for one digit count each (parallel compilation).
#pragma once
#include <type_traits>

#include "common.h"

typedef int v4i __attribute__((ext_vector_type(4)));

// K slicing of one launch: S slices of kslice each; slice s writes at
// C + s*cslice
struct Slices {
    int S = 1;
    int64_t kslice = 0;
    int64_t cslice = 0;
};

struct LimbW {
    double w[7];  // 256^u mod p
};

// Launch the limb core for L digits on planes Al (Mp x Kp per digit) and Bl
// (Np x Kp per digit) with a tile_m x tile_n output tile; defined for
// L = 1..4 in gemm_limb_l<L>.hip.
template <int L>
int launch_limb_impl(modp_ctx *ctx, const GemmArgs &g, int epi, const int8_t *Al, int64_t Mp, const int8_t *Bl,
                     int64_t Np, int64_t Kp, const LimbW &lw, int tile_m, int tile_n, const Slices &sl);

namespace {

constexpr int LBK = 64;  // k per MFMA / LDS stage

// Balanced base-256 digits of a residue.  top = 127*(256^L-1)/255: residues
// above it are taken as r - p, which lies in [-128*(256^L-1)/255, -1] because
// p <= 256^L, so L signed digits represent it exactly.
__device__ __forceinline__ int32_t digit_rep(uint32_t r, int32_t top, const ModP &mp)
{
    int32_t x = (int32_t)r;
    return x > top ? x - (int32_t)mp.p32 : x;
}
__device__ __forceinline__ int32_t next_digit(int32_t &x)
{
    const int32_t d = (int32_t)(int8_t)(x & 255);
    x = (x - d) >> 8;
    return d;
}

// A (m x k col-major) -> Al[s][row][kk] (row-major per digit, k contiguous),
// zero-padded to Mp x Kp.  64x64 tiles transposed through LDS.
template <typename TA>
__global__ __launch_bounds__(256) void k_split_a(int64_t m, int64_t k, const TA *__restrict__ A, int64_t lda,
                                                 int8_t *__restrict__ Al, int64_t Mp, int64_t Kp, int L,
                                                 int32_t top, ModP mp, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ int32_t tile[64][65];
    const int tid = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.y * 64, k0 = (int64_t)blockIdx.x * 64;
    if (r0 + 64 <= m && k0 + 64 <= k) {
        // a tile wholly inside A (the same for the whole workgroup): the 16
        // loads of a thread go out together
        const int row = tid & 63;
        TA x[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = A[r0 + row + (k0 + (tid >> 6) + 4 * i) * lda];
#pragma unroll
        for (int i = 0; i < 16; ++i) tile[row][(tid >> 6) + 4 * i] = digit_rep(to_res<TA>(x[i], mp), top, mp);
    } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            int row = tid & 63, kk = (tid >> 6) + 4 * i;
            int64_t gr = r0 + row, gk = k0 + kk;
            int32_t v = 0;
            if (gr < m && gk < k) v = digit_rep(to_res<TA>(A[gr + gk * lda], mp), top, mp);
            tile[row][kk] = v;
        }
    }
    __syncthreads();
    const int row = tid >> 2, ch = tid & 3;
    const size_t plane = (size_t)Mp * Kp;
    int8_t *dst = Al + (size_t)(r0 + row) * Kp + k0 + ch * 16;
    union { int8_t b[16]; v4i v; } u[4];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        int32_t x = tile[row][ch * 16 + e];
#pragma unroll
        for (int s = 0; s < 4; ++s) u[s].b[e] = (int8_t)next_digit(x);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
        if (s < L) *reinterpret_cast<v4i *>(dst + s * plane) = u[s].v;
}

// B (k x n col-major) -> Bl[s][col][kk], zero-padded to Np x Kp.  Same
// orientation (k contiguous), so each thread maps 4 consecutive k of one
// column: consecutive threads read consecutive 32-byte runs.
template <typename TB>
__global__ __launch_bounds__(256) void k_split_b(int64_t k, int64_t n, const TB *__restrict__ B, int64_t ldb,
                                                 int8_t *__restrict__ Bl, int64_t Np, int64_t Kp, int L,
                                                 int32_t top, ModP mp, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const int64_t quads = Kp / 4;
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= Np * quads) return;
    const int64_t col = e / quads, k0 = (e % quads) * 4;
    uint32_t word[4] = {0u, 0u, 0u, 0u};
    int32_t xs[4];
    if (col < n && k0 + 4 <= k) {
        // all four in range: the loads go out together
        TB v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = B[k0 + i + col * ldb];
#pragma unroll
        for (int i = 0; i < 4; ++i) xs[i] = digit_rep(to_res<TB>(v[i], mp), top, mp);
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t gk = k0 + i;
            xs[i] = (col < n && gk < k) ? digit_rep(to_res<TB>(B[gk + col * ldb], mp), top, mp) : 0;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int32_t x = xs[i];
#pragma unroll
        for (int s = 0; s < 4; ++s) word[s] |= ((uint32_t)next_digit(x) & 255u) << (8 * i);
    }
    const size_t plane = (size_t)Np * Kp;
    int8_t *dst = Bl + (size_t)col * Kp + k0;
#pragma unroll
    for (int s = 0; s < 4; ++s)
        if (s < L) *reinterpret_cast<uint32_t *>(dst + s * plane) = word[s];
}

// Limb-form tile shapes: a BM x BN output tile per workgroup of WGM x WGN
// waves, each wave a (BM/WGM) x (BN/WGN) block of MF x MF MFMA blocks: MF = 32
// uses v_mfma_i32_32x32x32_i8 (two k-steps per 64-deep stage), MF = 16 uses
// v_mfma_i32_16x16x64_i8.  On gfx950 the 32x32 form sustains about 1.4x the
// int8 rate of the 16x16 form (3.96 vs 2.8 POP/s measured).
// LDS rows hold the 64 digits of one row (A) or column (B) for one K stage
// as four 16-byte chunks.  SWZ = 0: rows padded to 80 bytes.  SWZ = 1: rows of
// 64 bytes with chunk q of row r stored at slot q ^ ((r >> 2) & 3), so both
// the 16-lane row-wise fragment reads and the 4-lanes-per-row stores touch
// every bank group once.  DB = 1: two LDS stages; stage k+1 is stored while
// stage k feeds the MFMAs, with one barrier per stage.  DB = 2: in addition
// two register staging sets, so each global load has two stages to land.
template <int BM_, int BN_, int WGM_, int WGN_, int MF_ = 16, int SWZ_ = 1, int DB_ = 0, int OCC_ = 1>
struct LimbShape {
    static constexpr int BM = BM_, BN = BN_, WGM = WGM_, WGN = WGN_, MF = MF_, SWZ = SWZ_, DB = DB_;
    static constexpr int OCC = OCC_;  // waves per SIMD the register budget must allow
    static constexpr int NT = 64 * WGM * WGN;
    static constexpr int PITCH = SWZ ? 64 : 80;
};

typedef int v16i __attribute__((ext_vector_type(16)));

template <int SWZ> __device__ __forceinline__ int lds_slot(int row, int q)
{
    return SWZ ? (q ^ ((row >> 2) & 3)) : q;
}

template <int L, typename SH, typename TC, int EPI>
__global__ __launch_bounds__(SH::NT, SH::OCC) void k_gemm_limb(int64_t m, int64_t n, int64_t Kp,
                                                     const int8_t *__restrict__ Al, int64_t Mp,
                                                     const int8_t *__restrict__ Bl, int64_t Np,
                                                     TC *__restrict__ C, int64_t ldc, ModP mp, LimbW lw,
                                                     int64_t kslice, int64_t cslice,
                                                     const int *__restrict__ skip)
{
    if (skip && *skip) return;
    constexpr int BM = SH::BM, BN = SH::BN, NT = SH::NT, PITCH = SH::PITCH, SWZ = SH::SWZ;
    constexpr int NS = 2 * L - 1;
    constexpr int MF = SH::MF;
    constexpr int WM = BM / SH::WGM, WN = BN / SH::WGN;
    constexpr int FM = WM / MF, FN = WN / MF;
    constexpr int KS = MF == 32 ? 2 : 1;     // MFMA k-steps per 64-deep stage
    constexpr int AE = MF == 32 ? 16 : 4;    // accumulator elements per lane
    typedef std::conditional_t<MF == 32, v16i, v4i> acc_t;
    constexpr int ACH = L * BM * (LBK / 16);  // 16-byte chunks of A per stage
    constexpr int BCH = L * BN * (LBK / 16);
    constexpr int NBUF = SH::DB ? 2 : 1;
    __shared__ __attribute__((aligned(16))) int8_t As[NBUF][L][BM][PITCH];
    __shared__ __attribute__((aligned(16))) int8_t Bs[NBUF][L][BN][PITCH];

    int64_t tm, tn;
    if (!tile_of_block(Mp / BM, Np / BN, &tm, &tn)) return;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % SH::WGM, wn = wave / SH::WGM;
    const int64_t r0 = tm * BM;
    const int64_t c0 = tn * BN;
    const size_t aplane = (size_t)Mp * Kp, bplane = (size_t)Np * Kp;
    // K slice of this workgroup (whole K when kslice >= Kp)
    const int64_t kbeg = (int64_t)blockIdx.y * kslice;
    const int64_t kend = Kp - kbeg < kslice ? Kp : kbeg + kslice;
    C += blockIdx.y * cslice;

    acc_t acc[FM][FN][NS];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int u = 0; u < NS; ++u) acc[i][j][u] = acc_t{};

    constexpr int AR = (ACH + NT - 1) / NT, BR = (BCH + NT - 1) / NT;
    constexpr int NSET = SH::DB == 2 ? 2 : 1;
    v4i ra[NSET][AR], rb[NSET][BR];
    auto load = [&](int set, int64_t k0) {
#pragma unroll
        for (int i = 0; i < AR; ++i) {
            int idx = tid + NT * i;
            if (ACH % NT == 0 || idx < ACH) {
                int s = idx / (BM * 4), rem = idx % (BM * 4);
                int row = rem >> 2, ch = rem & 3;
                ra[set][i] = *reinterpret_cast<const v4i *>(Al + s * aplane + (size_t)(r0 + row) * Kp + k0 + ch * 16);
            }
        }
#pragma unroll
        for (int i = 0; i < BR; ++i) {
            int idx = tid + NT * i;
            if (BCH % NT == 0 || idx < BCH) {
                int s = idx / (BN * 4), rem = idx % (BN * 4);
                int col = rem >> 2, ch = rem & 3;
                rb[set][i] = *reinterpret_cast<const v4i *>(Bl + s * bplane + (size_t)(c0 + col) * Kp + k0 + ch * 16);
            }
        }
    };
    auto stash = [&](int set, int buf) {
#pragma unroll
        for (int i = 0; i < AR; ++i) {
            int idx = tid + NT * i;
            if (ACH % NT == 0 || idx < ACH) {
                int s = idx / (BM * 4), rem = idx % (BM * 4);
                int row = rem >> 2;
                *reinterpret_cast<v4i *>(&As[buf][s][row][lds_slot<SWZ>(row, rem & 3) * 16]) = ra[set][i];
            }
        }
#pragma unroll
        for (int i = 0; i < BR; ++i) {
            int idx = tid + NT * i;
            if (BCH % NT == 0 || idx < BCH) {
                int s = idx / (BN * 4), rem = idx % (BN * 4);
                int col = rem >> 2;
                *reinterpret_cast<v4i *>(&Bs[buf][s][col][lds_slot<SWZ>(col, rem & 3) * 16]) = rb[set][i];
            }
        }
    };
    // fragments: lane holds row (lane % MF) of the block and 16 digits of
    // k-chunk q (16x16x64: q = lane / 16; 32x32x32: q = 2 ks + lane / 32)
    v4i a[FM][KS][L], b[FN][KS][L];
    auto fragments = [&](int buf) {
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int q = MF == 32 ? ks * 2 + (lane >> 5) : (lane >> 4);
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                const int row = wm * WM + i * MF + (lane & (MF - 1));
#pragma unroll
                for (int s = 0; s < L; ++s)
                    a[i][ks][s] = *reinterpret_cast<const v4i *>(&As[buf][s][row][lds_slot<SWZ>(row, q) * 16]);
            }
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int col = wn * WN + j * MF + (lane & (MF - 1));
#pragma unroll
                for (int t = 0; t < L; ++t)
                    b[j][ks][t] = *reinterpret_cast<const v4i *>(&Bs[buf][t][col][lds_slot<SWZ>(col, q) * 16]);
            }
        }
    };
    auto mfmas = [&]() {
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int s = 0; s < L; ++s)
#pragma unroll
                        for (int t = 0; t < L; ++t) {
                            if constexpr (MF == 32)
                                acc[i][j][s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[i][ks][s], b[j][ks][t],
                                                                                         acc[i][j][s + t], 0, 0, 0);
                            else
                                acc[i][j][s + t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[i][ks][s], b[j][ks][t],
                                                                                         acc[i][j][s + t], 0, 0, 0);
                        }
    };

    const int64_t nk = (kend - kbeg) / LBK;
    if constexpr (SH::DB == 2) {
        // stage kt lives in LDS buffer kt & 1; register set (kt + 1) & 1
        // holds stage kt + 1 and set kt & 1 receives stage kt + 2
        load(0, kbeg);
        if (nk > 1) load(1, kbeg + LBK);
        stash(0, 0);
        if (nk > 2) load(0, kbeg + 2 * LBK);
        __syncthreads();
        auto step = [&](int64_t kt, auto cur_c) {
            constexpr int cur = decltype(cur_c)::value;
            fragments(cur);
            if (kt + 1 < nk) stash(cur ^ 1, cur ^ 1);
            if (kt + 3 < nk) load(cur ^ 1, kbeg + (kt + 3) * LBK);
            mfmas();
            __syncthreads();
        };
        int64_t kt = 0;
        for (; kt + 1 < nk; kt += 2) {
            step(kt, std::integral_constant<int, 0>());
            step(kt + 1, std::integral_constant<int, 1>());
        }
        if (kt < nk) step(kt, std::integral_constant<int, 0>());
    } else if constexpr (SH::DB == 1) {
        // stage kt lives in buffer kt & 1; the registers hold stage kt + 1
        load(0, kbeg);
        stash(0, 0);
        if (nk > 1) load(0, kbeg + LBK);
        __syncthreads();
        for (int64_t kt = 0; kt < nk; ++kt) {
            const int cur = (int)(kt & 1);
            fragments(cur);
            // buffer cur ^ 1 was last read before the previous barrier
            if (kt + 1 < nk) stash(0, cur ^ 1);
            if (kt + 2 < nk) load(0, kbeg + (kt + 2) * LBK);
            mfmas();
            __syncthreads();
        }
    } else {
        load(0, kbeg);
        for (int64_t kt = 0; kt < nk; ++kt) {
            stash(0, 0);
            __syncthreads();
            if (kt + 1 < nk) load(0, kbeg + (kt + 1) * LBK);
            fragments(0);
            mfmas();
            __syncthreads();
        }
    }

    // accumulator layouts: 16x16 blocks hold column lane % 16, rows
    // 4 (lane / 16) + e; 32x32 blocks hold column lane % 32, rows
    // 8 (e / 4) + 4 (lane / 32) + e % 4
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            int64_t gc = c0 + wn * WN + j * MF + (lane & (MF - 1));
            if (gc >= n) continue;
#pragma unroll
            for (int e = 0; e < AE; ++e) {
                const int rr = MF == 32 ? 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3) : 4 * (lane >> 4) + e;
                int64_t gr = r0 + wm * WM + i * MF + rr;
                if (gr >= m) continue;
                double tot = dmod((double)acc[i][j][0][e], mp);
#pragma unroll
                for (int u = 1; u < NS; ++u) {
                    double d = dmod((double)acc[i][j][u][e], mp);
                    tot = dmod(fma(d, lw.w[u], tot), mp);
                }
                uint32_t r = (uint32_t)tot;
                TC *dst = C + gr + gc * ldc;
                if (EPI == EPI_SUB) r = submod(to_res<TC>(*dst, mp), r, mp);
                else if (EPI == EPI_ADD) r = addmod(to_res<TC>(*dst, mp), r, mp);
                *dst = from_res<TC>(r);
            }
        }
}


// MFMA block size and LDS layout used by the library's limb kernels
constexpr int LIMB_MF = 16;
constexpr int LIMB_SWZ = 1;

template <int L, typename SH, typename TC>
int limb_core(modp_ctx *ctx, const GemmArgs &g, int epi, const int8_t *Al, int64_t Mp, const int8_t *Bl, int64_t Np,
              int64_t Kp, const LimbW &lw, const Slices &sl)
{
    const dim3 grid(xcd_grid((Np / SH::BN) * (Mp / SH::BM)), (unsigned)sl.S);
    TimedLaunch tl(ctx, "gemm_limb");
    TC *C = (TC *)g.C;
    if (epi == EPI_STORE)
        k_gemm_limb<L, SH, TC, EPI_STORE><<<grid, SH::NT, 0, ctx->stream>>>(
            g.m, g.n, Kp, Al, Mp, Bl, Np, C, g.ldc, g.mp, lw, sl.kslice, sl.cslice, g.skip);
    else if (epi == EPI_SUB)
        k_gemm_limb<L, SH, TC, EPI_SUB><<<grid, SH::NT, 0, ctx->stream>>>(
            g.m, g.n, Kp, Al, Mp, Bl, Np, C, g.ldc, g.mp, lw, sl.kslice, sl.cslice, g.skip);
    else
        k_gemm_limb<L, SH, TC, EPI_ADD><<<grid, SH::NT, 0, ctx->stream>>>(
            g.m, g.n, Kp, Al, Mp, Bl, Np, C, g.ldc, g.mp, lw, sl.kslice, sl.cslice, g.skip);
    return hip_check(ctx, hipGetLastError(), "k_gemm_limb launch");
}

// tile_m x tile_n -> workgroup shape: 64x64 and 128x64 use 32x32 waves;
// 128x128 (L <= 2 only; the caller pads to the tile it asks for) uses 2x2
// waves of 64x64
template <int L, typename TC>
int limb_by_shape(modp_ctx *ctx, const GemmArgs &g, int epi, const int8_t *Al, int64_t Mp, const int8_t *Bl,
                  int64_t Np, int64_t Kp, const LimbW &lw, int tile_m, int tile_n, const Slices &sl)
{
    if (L <= 2 && tile_m == 128 && tile_n == 128)
        return limb_core<L, LimbShape<128, 128, 2, 2, LIMB_MF, LIMB_SWZ>, TC>(ctx, g, epi, Al, Mp, Bl, Np, Kp, lw, sl);
    if (tile_m == 128 && tile_n == 64)
        return limb_core<L, LimbShape<128, 64, 4, 2, LIMB_MF, LIMB_SWZ>, TC>(ctx, g, epi, Al, Mp, Bl, Np, Kp, lw, sl);
    return limb_core<L, LimbShape<64, 64, 2, 2, LIMB_MF, LIMB_SWZ>, TC>(ctx, g, epi, Al, Mp, Bl, Np, Kp, lw, sl);
}

}  // namespace

template <int L>
int launch_limb_impl(modp_ctx *ctx, const GemmArgs &g, int epi, const int8_t *Al, int64_t Mp, const int8_t *Bl,
                     int64_t Np, int64_t Kp, const LimbW &lw, int tile_m, int tile_n, const Slices &sl)
{
    if (g.kc == EK_F64) return limb_by_shape<L, double>(ctx, g, epi, Al, Mp, Bl, Np, Kp, lw, tile_m, tile_n, sl);
    if (g.kc == EK_F32) return limb_by_shape<L, float>(ctx, g, epi, Al, Mp, Bl, Np, Kp, lw, tile_m, tile_n, sl);
    return limb_by_shape<L, int32_t>(ctx, g, epi, Al, Mp, Bl, Np, Kp, lw, tile_m, tile_n, sl);
}
