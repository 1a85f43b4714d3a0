// gemm_rns.h -- multi-modular (residue number system) form of the mod-p GEMM
// on gfx950 int8 MFMA.
//
// C = A*B mod p is computed through the exact integer product of the
// balanced representatives (|a|, |b| <= floor(p/2)):  C_int = sum_k a b,
// |C_int| <= k floor(p/2)^2.  That integer is taken modulo r pairwise coprime
// moduli q_0..q_{r-1} <= 256 with M = prod q_j > 2 |C_int|:
//
//   * split: every operand entry is reduced mod each q_j into a residue
//     held in int8 (one plane per modulus, |residue| <= 128); the B side is
//     multiplied by y_j = (M/q_j)^{-1} mod q_j on the way;
//   * core: one int8 GEMM per modulus on v_mfma_i32_32x32x32_i8 with int32
//     accumulation (|partial sums| <= k 2^14 < 2^31 for k <= 131071), whose
//     epilogue writes e_j = D_j mod q_j = (y_j C_int) mod q_j as one byte;
//   * combine: C_int = sum_j e_j M/q_j - t M with t = round(sum_j e_j / q_j)
//     (exact: |C_int| / M < 1/2 leaves a margin far above the FP64 error of
//     the sum), so C mod p = (sum_j e_j (M/q_j mod p) - t (M mod p)) mod p,
//     every term an exact integer below 2^40 in FP64.
//
// This is the Chinese-remainder reconstruction (Garner 1959; Knuth TAOCP
// vol. 2 4.3.2) applied to an integer matrix product, as in the Ozaki
// scheme II for emulating FP64 GEMM with int8 matrix units (Ozaki, Uchino,
// Imamura 2025).  For p close to 2^26 and k = 8192 it needs 9 int8 products
// where the base-256 digit form needs 16.
#pragma once
#include "common.h"

namespace rns {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int NMOD_MAX = 12;
// pairwise coprime moduli <= 256, largest first:
// 2^8, 3*5*17, 11*23, 251, 13*19, 241, 239, 233, 229, 227, 223, 211
constexpr int QLIST[NMOD_MAX] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 211};

// per-launch constants of the r moduli (kernel argument)
struct Moduli {
    int r;
    int single;               // one plane with q_0 = p (2 <= p <= 255): C mod p directly
    int q[NMOD_MAX];
    int y[NMOD_MAX];          // (M/q_j)^{-1} mod q_j
    double w[NMOD_MAX];       // (M/q_j) mod p
    double invq[NMOD_MAX];    // 1/q_j
    double wm;                // M mod p
};

// A representative of t mod Q in int8 for an integer |t| < 3.2e6 held
// exactly in a float: r = t - Q rint(t / Q) with the quotient taken from the
// rounded product t (1/Q).  That product is within 2^-9 of t / Q, and for odd
// Q the fraction t / Q - n is never closer than 1 / (2Q) > 2^-9 to 1/2, so r
// is the balanced residue, |r| <= (Q - 1) / 2 <= 127; for Q = 256, r lies in
// [-128, 128] and the byte cast maps 128 to -128 (the same residue mod 256).
// The fma is exact (every term is an integer below 2^24).
template <int Q> __device__ __forceinline__ int res8(float t)
{
    constexpr float QF = (float)Q, IQ = 1.0f / (float)Q;
    return (int)fmaf(-rintf(t * IQ), QF, t);
}

// the 16 residue bytes of x[0..15] mod Q, packed little-endian, with
// x = xh 4096 + xl (|xh| <= 2^13, 0 <= xl < 4096):
//   A side  t = xh (4096 mod Q) + xl,          |t| < 2^21;
//   B side  t = xh (4096 y mod Q) + xl y,      |t| < 255 (2^13 + 4096),
// i.e. x y mod Q in one reduction (y < Q is the CRT weight of the plane).
template <int Q, bool mulY> __device__ __forceinline__ v4i pack16q(const float *xh, const float *xl, int y)
{
    const float C = mulY ? (float)((4096 * y) % Q) : (float)(4096 % Q);
    const float yf = (float)y;
    v4i w;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        uint32_t u = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float lo = mulY ? xl[4 * d + e] * yf : xl[4 * d + e];
            const int v = res8<Q>(fmaf(xh[4 * d + e], C, lo));
            u |= ((uint32_t)v & 255u) << (8 * e);
        }
        w[d] = (int)u;
    }
    return w;
}

// pack16q for a modulus q known at run time (the single-plane form, q = p
// <= 255).  There |x| <= 127, so |t| < 4096 + 255 and t (1/q) is within
// 2^-20 of t / q, far closer than a non-tie quotient is to a half-integer
// (1 / (2q) for odd q, 1 / q for even q): r is the balanced residue, or +-q/2
// at an exact tie of an even q; |r| <= 127 either way.
template <bool mulY> __device__ __forceinline__ v4i pack16r(const float *xh, const float *xl, int q, int y)
{
    const float qf = (float)q, iq = 1.0f / qf;
    const float C = mulY ? (float)((4096 * y) % q) : (float)(4096 % q);
    const float yf = (float)y;
    v4i w;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        uint32_t u = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float lo = mulY ? xl[4 * d + e] * yf : xl[4 * d + e];
            const float t = fmaf(xh[4 * d + e], C, lo);
            const int v = (int)fmaf(-rintf(t * iq), qf, t);
            u |= ((uint32_t)v & 255u) << (8 * e);
        }
        w[d] = (int)u;
    }
    return w;
}

// x -> (xh, xl) as floats: x = xh 4096 + xl with 0 <= xl < 4096
__device__ __forceinline__ void split12(int x, float &xh, float &xl)
{
    xh = (float)(x >> 12);
    xl = (float)(x & 4095);
}

// the same for modulus q_j (a wave-uniform j: one branch per plane)
template <bool mulY> __device__ __forceinline__ v4i pack16(const float *xh, const float *xl, int j, const Moduli &md)
{
    if (md.single) return pack16r<mulY>(xh, xl, md.q[0], md.y[0]);
    switch (j) {
    case 0: return pack16q<QLIST[0], mulY>(xh, xl, md.y[0]);
    case 1: return pack16q<QLIST[1], mulY>(xh, xl, md.y[1]);
    case 2: return pack16q<QLIST[2], mulY>(xh, xl, md.y[2]);
    case 3: return pack16q<QLIST[3], mulY>(xh, xl, md.y[3]);
    case 4: return pack16q<QLIST[4], mulY>(xh, xl, md.y[4]);
    case 5: return pack16q<QLIST[5], mulY>(xh, xl, md.y[5]);
    case 6: return pack16q<QLIST[6], mulY>(xh, xl, md.y[6]);
    case 7: return pack16q<QLIST[7], mulY>(xh, xl, md.y[7]);
    case 8: return pack16q<QLIST[8], mulY>(xh, xl, md.y[8]);
    case 9: return pack16q<QLIST[9], mulY>(xh, xl, md.y[9]);
    case 10: return pack16q<QLIST[10], mulY>(xh, xl, md.y[10]);
    default: return pack16q<QLIST[11], mulY>(xh, xl, md.y[11]);
    }
}

// ---------------------------------------------------------------------------
// core: one int8 GEMM per modulus (blockIdx.y), D = X * Y^T with X the n side
// (planes [r][Np][Kp], row = a column of B) and Y the m side (planes
// [r][Mp][Kp], row = a row of A); D[i][j] is C(j, i), written as the residue
// byte e = D mod q at E[plane][i * Mp + j] (C's column-major layout, ld Mp).
// ---------------------------------------------------------------------------

// TN x TM output tile (n side x m side) per workgroup of WGN x WGM waves; BK
// k per LDS stage (two stages, one barrier per stage)
template <int TN_, int TM_, int WGN_, int WGM_, int BK_>
struct Shape {
    static constexpr int TN = TN_, TM = TM_, WGN = WGN_, WGM = WGM_, BK = BK_;
    static constexpr int NT = 64 * WGN * WGM;
    static constexpr int WTN = TN / WGN, WTM = TM / WGM;  // wave tile
    static constexpr int FN = WTN / 32, FM = WTM / 32;    // 32x32 blocks per wave
    static constexpr int CPR = BK / 16;                   // 16-byte chunks per LDS row
    static constexpr int KK = BK / 32;                    // MFMA k-steps per stage
};

// LDS rows of BK = 128 bytes hold 8 16-byte chunks; chunk q of row `row`
// sits in slot q ^ ((row >> 1) & 7).  The LDS has 64 banks of 4 bytes (256
// bytes per bank row), so rows r and r + 1 fall in opposite halves of the
// banks and the 16-byte bank group of (row, slot) is 8 (row & 1) + slot: the
// 16 consecutive rows of one fragment read at one chunk (16 lanes per LDS
// cycle) and the 2 rows of 8 lanes of a staging write all land in distinct
// bank groups (the counters show no bank conflicts; q ^ (row & 7) left
// rows r and r + 8 on one group).
template <int CPR> __device__ __forceinline__ int slot(int row, int q)
{
    static_assert(CPR == 8, "128-byte LDS rows");
    return q ^ ((row >> 1) & 7);
}

// The core with an explicitly scheduled main loop.  Per thread, every
// global chunk of a stage sits at one 32-bit offset from a workgroup-uniform
// base (rows 32 apart add a uniform step) and every LDS chunk at one offset
// plus an immediate; the fragments of k-step kk + 1 are read while the
// MFMAs of k-step kk run (two register sets), the staged registers of stage
// s + 1 are written to LDS during the first two k-steps and stage s + 2 is
// loaded during the last two, each group interleaved with the MFMAs by
// scheduling barriers.  Steady-state stages are branch free (the last two
// are peeled).
template <class SH>
__global__ __launch_bounds__(SH::NT, 1) void k_rns_gemm2(const int8_t *__restrict__ X, const int8_t *__restrict__ Y,
                                                       int64_t Np, int64_t Mp, int64_t Kp,
                                                       uint8_t *__restrict__ E, Moduli md)
{
    constexpr int TN = SH::TN, TM = SH::TM, BK = SH::BK, NT = SH::NT, CPR = SH::CPR, KK = SH::KK;
    constexpr int FN = SH::FN, FM = SH::FM;
    constexpr int XC = TN * CPR / NT, YC = TM * CPR / NT;
    constexpr int RSTEP = NT / CPR;  // rows between a thread's consecutive chunks
    static_assert(CPR == 8 && KK == 4, "this schedule is written for 128-deep stages");
    static_assert(XC % 2 == 0 && YC % 2 == 0, "stage chunks are written and loaded in halves");
    __shared__ __attribute__((aligned(16))) int8_t Xs[2][TN * BK];
    __shared__ __attribute__((aligned(16))) int8_t Ys[2][TM * BK];

    int64_t ti, tj;
    if (!tile_of_block(Np / TN, Mp / TM, &ti, &tj)) return;
    const int plane = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wn = wave % SH::WGN, wm = wave / SH::WGN;
    const int64_t n0 = ti * TN, m0 = tj * TM;
    const int8_t *Xb = X + (size_t)plane * Np * Kp + (size_t)n0 * Kp;
    const int8_t *Yb = Y + (size_t)plane * Mp * Kp + (size_t)m0 * Kp;
    const int64_t rstep = (int64_t)RSTEP * Kp;

    // staging: chunk c of this thread is row tid / CPR + RSTEP c, chunk tid % CPR
    const int srow = tid / CPR, sq = tid % CPR;
    const uint32_t goff = (uint32_t)(srow * Kp + 16 * sq);
    const uint32_t loff = (uint32_t)(srow * BK + 16 * slot<CPR>(srow, sq));
    // fragments: rows (w * F + a) * 32 + lane % 32 all share (row >> 1) % 8
    const int h = lane >> 5, l7 = (lane >> 1) & 7;
    uint32_t fx[KK], fy[KK];
#pragma unroll
    for (int kk = 0; kk < KK; ++kk) {
        const int sl = (2 * kk + h) ^ l7;
        fx[kk] = (uint32_t)((wn * FN * 32 + (lane & 31)) * BK + 16 * sl);
        fy[kk] = (uint32_t)((wm * FM * 32 + (lane & 31)) * BK + 16 * sl);
    }

    v16i acc[FN][FM];
#pragma unroll
    for (int a = 0; a < FN; ++a)
#pragma unroll
        for (int b = 0; b < FM; ++b) acc[a][b] = v16i{0};

    v4i rx[XC], ry[YC];
    v4i xf[2][FN], yf[2][FM];

    auto gload_half = [&](int64_t k0, int half) {
#pragma unroll
        for (int c = half * XC / 2; c < (half + 1) * XC / 2; ++c)
            rx[c] = *reinterpret_cast<const v4i *>(Xb + c * rstep + k0 + goff);
#pragma unroll
        for (int c = half * YC / 2; c < (half + 1) * YC / 2; ++c)
            ry[c] = *reinterpret_cast<const v4i *>(Yb + c * rstep + k0 + goff);
    };
    auto lstore_half = [&](int buf, int half) {
#pragma unroll
        for (int c = half * XC / 2; c < (half + 1) * XC / 2; ++c)
            *reinterpret_cast<v4i *>(&Xs[buf][loff + c * RSTEP * BK]) = rx[c];
#pragma unroll
        for (int c = half * YC / 2; c < (half + 1) * YC / 2; ++c)
            *reinterpret_cast<v4i *>(&Ys[buf][loff + c * RSTEP * BK]) = ry[c];
    };
    auto lfrag = [&](int buf, int kk, int set) {
#pragma unroll
        for (int a = 0; a < FN; ++a) xf[set][a] = *reinterpret_cast<const v4i *>(&Xs[buf][fx[kk] + a * 32 * BK]);
#pragma unroll
        for (int b = 0; b < FM; ++b) yf[set][b] = *reinterpret_cast<const v4i *>(&Ys[buf][fy[kk] + b * 32 * BK]);
    };
    auto mfmas = [&](int set) {
#pragma unroll
        for (int a = 0; a < FN; ++a)
#pragma unroll
            for (int b = 0; b < FM; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(yf[set][b], xf[set][a], acc[a][b], 0, 0, 0);
    };
    constexpr int NMF = FN * FM, NFR = FN + FM, NST = (XC + YC) / 2, NGL = (XC + YC) / 2;
    // one k-step: its MFMAs interleaved with the fragment reads of the next
    // k-step and (when asked) half of the stage's LDS writes or global loads
    auto step = [&](int buf, int kk, bool st_half, int sthalf, int nbuf, bool gl_half, int glhalf, int64_t k2) {
        const int set = kk & 1;
        if (kk + 1 < KK) lfrag(buf, kk + 1, set ^ 1);
        if (st_half) lstore_half(nbuf, sthalf);
        if (gl_half) gload_half(k2, glhalf);
        mfmas(set);
        // the pattern: an MFMA, then one memory instruction, repeated
#pragma unroll
        for (int t = 0; t < NMF; ++t) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            if (kk + 1 < KK && t < NFR) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
            if (st_half && t < NST) __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
            if (gl_half && t < NGL) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
        }
    };

    const int64_t ns = Kp / BK;
    gload_half(0, 0);
    gload_half(0, 1);
    lstore_half(0, 0);
    lstore_half(0, 1);
    if (ns > 1) {
        gload_half(BK, 0);
        gload_half(BK, 1);
    }
    __syncthreads();
    lfrag(0, 0, 0);
    int64_t s = 0;
    // steady state: stage s stores stage s + 1 and loads stage s + 2
    for (; s + 2 < ns; ++s) {
        const int buf = (int)(s & 1);
        const int64_t k2 = (s + 2) * BK;
        step(buf, 0, true, 0, buf ^ 1, false, 0, 0);
        step(buf, 1, true, 1, buf ^ 1, false, 0, 0);
        step(buf, 2, false, 0, 0, true, 0, k2);
        step(buf, 3, false, 0, 0, true, 1, k2);
        __syncthreads();
        lfrag(buf ^ 1, 0, 0);
    }
    if (s + 1 < ns) {
        const int buf = (int)(s & 1);
        step(buf, 0, true, 0, buf ^ 1, false, 0, 0);
        step(buf, 1, true, 1, buf ^ 1, false, 0, 0);
        step(buf, 2, false, 0, 0, false, 0, 0);
        step(buf, 3, false, 0, 0, false, 0, 0);
        __syncthreads();
        lfrag(buf ^ 1, 0, 0);
        ++s;
    }
    {
        const int buf = (int)(s & 1);
#pragma unroll
        for (int kk = 0; kk < KK; ++kk) step(buf, kk, false, 0, 0, false, 0, 0);
    }

    // Epilogue.  The blocks are D^T (Y rows first), so lane l holds E row
    // i = column l % 32 of the block and the four consecutive j = 8 g + 4 (l / 32)
    // + 0..3 of group g, packed into one word.  v mod q in FP32: v = a 2^16 + b
    // (|a| <= 2^15, 0 <= b < 2^16), t = a (2^16 mod q) + b is exact and below
    // 2^23, rint(t / q) is within 0.01 of t / q, so r = t - q rint(t / q) is
    // exact with |r| < q and one fix-up gives v mod q in [0, q).
    // The words go through LDS (the stage buffer that the last stage does not
    // read: the barrier closing stage ns - 2 released it) so that every store
    // writes 16 whole 64-byte row segments: the wave's WTN x WTM bytes sit in
    // rows of WTM bytes whose 16-byte chunk c is stored at c ^ ((row / 4) % 4)
    // (the word writes of 32 rows spread over 32 banks, the 16-byte reads of 4
    // consecutive rows cover all 64).
    constexpr int WTN = SH::WTN, WTM = SH::WTM, NW = NT / 64;
    static_assert(WTM == 64 && NW % 2 == 0 && (NW / 2) * WTN * WTM <= TN * BK && TN == TM, "epilogue staging");
    const int q = md.q[plane];
    const float qf = (float)q, iqf = 1.0f / qf, c16 = (float)(65536 % q);
    const int nb = (int)(ns & 1);
    int8_t *stg = (wave < NW / 2 ? Xs[nb] : Ys[nb]) + (wave % (NW / 2)) * (WTN * WTM);
#pragma unroll
    for (int a = 0; a < FN; ++a) {
        const int il = a * 32 + (lane & 31);
#pragma unroll
        for (int b = 0; b < FM; ++b)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                uint32_t u = 0;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int v = acc[a][b][4 * g + t];
                    const float tv = fmaf((float)(v >> 16), c16, (float)(v & 0xFFFF));
                    int r = (int)fmaf(-rintf(tv * iqf), qf, tv);
                    r += r < 0 ? q : 0;
                    u |= (uint32_t)r << (8 * t);
                }
                const int chunk = 2 * b + (g >> 1), w = 2 * (g & 1) + (lane >> 5);
                *reinterpret_cast<uint32_t *>(stg + il * WTM + 16 * (chunk ^ ((il >> 2) & 3)) + 4 * w) = u;
            }
    }
    uint8_t *Eg = E + (size_t)plane * Np * Mp + (size_t)(n0 + wn * WTN) * Mp + m0 + wm * WTM;
#pragma unroll
    for (int it = 0; it < WTN / 16; ++it) {
        const int row = it * 16 + (lane >> 2), c = lane & 3;
        const v4i w4 = *reinterpret_cast<const v4i *>(stg + row * WTM + 16 * (c ^ ((row >> 2) & 3)));
        *reinterpret_cast<v4i *>(Eg + (size_t)row * Mp + 16 * c) = w4;
    }
}

}  // namespace rns

namespace rns {

// ---------------------------------------------------------------------------
// split: operand residues -> balanced int8 residue planes
// ---------------------------------------------------------------------------

// balanced representative of a storage element, |x| <= floor(p/2)
template <typename T> __device__ __forceinline__ int balanced(T v, const ModP &mp)
{
    const int x = (int)to_res<T>(v, mp);
    return x > mp.half ? x - (int)mp.p32 : x;
}

// A (m x k column-major) -> planes Y[j][row][kk] (k contiguous), zero-padded
// to Mp x Kp; 64 x 128 tiles (rows x k) transposed through LDS.  Loads: each
// lane reads two consecutive rows of one column (16 bytes for FP64), a wave
// two 512-byte column segments.  Stores: eight lanes write a row's 128 bytes
// of a plane, a wave eight full rows.  blockIdx.x: k tile, blockIdx.y: row
// tile.
template <typename TA>
__global__ __launch_bounds__(256) void k_split_rows(int64_t m, int64_t k, const TA *__restrict__ A, int64_t lda,
                                                   int8_t *__restrict__ Yp, int64_t Mp, int64_t Kp, ModP mp,
                                                   Moduli md, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    constexpr int TR = 64, TK = 128, LDT = TK + 1;  // odd pitch: conflict-free column writes
    __shared__ int32_t tile[TR * LDT];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t r0 = (int64_t)blockIdx.y * TR, k0 = (int64_t)blockIdx.x * TK;
    const int rr = 2 * (lane & 31);
    const int64_t gr = r0 + rr;
    // a tile wholly inside A with every row pair aligned (the same for the
    // whole workgroup): all sixteen pair loads are issued before the first
    // is used, instead of one memory round trip per load
    const bool interior = r0 + TR <= m && k0 + TK <= k && (lda & 1) == 0 &&
                          (((uintptr_t)A + (size_t)r0 * sizeof(TA)) & (2 * sizeof(TA) - 1)) == 0;
    if (interior) {
        TA v[TK / 8][2];
#pragma unroll
        for (int it = 0; it < TK / 8; ++it) {
            const int kk = 8 * it + 2 * wave + (lane >> 5);
            __builtin_memcpy(v[it], A + gr + (k0 + kk) * lda, sizeof(v[it]));
        }
#pragma unroll
        for (int it = 0; it < TK / 8; ++it) {
            const int kk = 8 * it + 2 * wave + (lane >> 5);
            tile[rr * LDT + kk] = balanced<TA>(v[it][0], mp);
            tile[(rr + 1) * LDT + kk] = balanced<TA>(v[it][1], mp);
        }
    } else {
        const bool pair = gr + 1 < m && (((uintptr_t)A + (size_t)gr * sizeof(TA)) & (2 * sizeof(TA) - 1)) == 0 &&
                          (lda & 1) == 0;
#pragma unroll 4
        for (int it = 0; it < TK / 8; ++it) {
            const int kk = 8 * it + 2 * wave + (lane >> 5);
            const int64_t gk = k0 + kk;
            int x0 = 0, x1 = 0;
            if (gk < k) {
                const TA *src = A + gr + gk * lda;
                if (pair) {
                    TA v2[2];
                    __builtin_memcpy(v2, src, sizeof(v2));
                    x0 = balanced<TA>(v2[0], mp);
                    x1 = balanced<TA>(v2[1], mp);
                } else {
                    if (gr < m) x0 = balanced<TA>(src[0], mp);
                    if (gr + 1 < m) x1 = balanced<TA>(src[1], mp);
                }
            }
            tile[rr * LDT + kk] = x0;
            tile[(rr + 1) * LDT + kk] = x1;
        }
    }
    __syncthreads();
    const size_t plane = (size_t)Mp * Kp;
    const int ch = tid & 7;
#pragma unroll 1
    for (int pass = 0; pass < 2; ++pass) {
        const int row = 32 * pass + (tid >> 3);
        float xh[16], xl[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) split12(tile[row * LDT + ch * 16 + e], xh[e], xl[e]);
        int8_t *dst = Yp + (size_t)(r0 + row) * Kp + k0 + ch * 16;
#pragma unroll 1
        for (int j = 0; j < md.r; ++j) *reinterpret_cast<v4i *>(dst + j * plane) = pack16<false>(xh, xl, j, md);
    }
}

// B (k x n column-major) -> planes X[j][col][kk] (k contiguous, times y_j),
// zero-padded to Np x Kp.  Same orientation: each thread maps 16 consecutive
// k of one column.
template <typename TB>
__global__ __launch_bounds__(256) void k_split_cols(int64_t k, int64_t n, const TB *__restrict__ B, int64_t ldb,
                                                   int8_t *__restrict__ Xp, int64_t Np, int64_t Kp, ModP mp,
                                                   Moduli md, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const int64_t chunks = Kp / 16;
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= Np * chunks) return;
    const int64_t col = e / chunks, k0 = (e % chunks) * 16;
    float xh[16], xl[16];
    const TB *src = B + k0 + col * ldb;
    if (col < n && k0 + 16 <= k && ((uintptr_t)src & 15) == 0) {
        // the whole chunk in range and 16-byte aligned: sizeof(TB) 16-byte loads
        v4i raw[sizeof(TB)];
#pragma unroll
        for (int i = 0; i < (int)sizeof(TB); ++i) raw[i] = reinterpret_cast<const v4i *>(src)[i];
        TB v[16];
        __builtin_memcpy(v, raw, sizeof(v));
#pragma unroll
        for (int t = 0; t < 16; ++t) split12(balanced<TB>(v[t], mp), xh[t], xl[t]);
    } else {
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const int64_t gk = k0 + t;
            split12((col < n && gk < k) ? balanced<TB>(src[t], mp) : 0, xh[t], xl[t]);
        }
    }
    const size_t plane = (size_t)Np * Kp;
    int8_t *dst = Xp + (size_t)col * Kp + k0;
#pragma unroll 1
    for (int j = 0; j < md.r; ++j) *reinterpret_cast<v4i *>(dst + j * plane) = pack16<true>(xh, xl, j, md);
}

// ---------------------------------------------------------------------------
// combine: the r residue bytes of each entry -> C = epi(C, C_int mod p)
// ---------------------------------------------------------------------------
template <typename TC, int EPI>
__global__ __launch_bounds__(256) void k_combine(int64_t m, int64_t n, const uint8_t *__restrict__ E, int64_t Mp,
                                                int64_t plane, Moduli md, TC *__restrict__ C, int64_t ldc,
                                                ModP mp, const int *__restrict__ skip, bool vec_ok)
{
    if (skip && *skip) return;
    // four consecutive rows of one column per thread: one 4-byte load per
    // modulus (Mp is a multiple of the output tile, so the load is aligned
    // and stays inside the plane)
    const int64_t mq = (m + 3) / 4;
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= mq * n) return;
    const int64_t j0 = 4 * (idx % mq), i = idx / mq;
    const uint8_t *src = E + i * Mp + j0;
    double S[4] = {0.0, 0.0, 0.0, 0.0}, T[4] = {0.0, 0.0, 0.0, 0.0};
    // every plane's word loaded before the first is used (planes past the
    // r-th re-read the last one, so no load waits behind a branch)
    uint32_t w[NMOD_MAX];
#pragma unroll
    for (int jj = 0; jj < NMOD_MAX; ++jj)
        w[jj] = *reinterpret_cast<const uint32_t *>(src + (size_t)(jj < md.r ? jj : md.r - 1) * plane);
#pragma unroll
    for (int jj = 0; jj < NMOD_MAX; ++jj) {
        if (jj < md.r) {
            const uint32_t w4 = w[jj];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const double v = (double)((w4 >> (8 * e)) & 255u);
                S[e] = fma(v, md.w[jj], S[e]);  // < 12 * 2^8 * 2^26: exact
                T[e] = fma(v, md.invq[jj], T[e]);
            }
        }
    }
    TC *dst = C + j0 + i * ldc;
    uint32_t r[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        r[e] = (uint32_t)dmod(fma(-rint(T[e]), md.wm, S[e]), mp);
        if (EPI != EPI_STORE && j0 + e < m) {
            const uint32_t c = to_res<TC>(dst[e], mp);
            r[e] = EPI == EPI_SUB ? submod(c, r[e], mp) : addmod(c, r[e], mp);
        }
    }
    // four whole rows at a 16-byte aligned address (float64, ldc even and C
    // aligned, which the launch checks): two 16-byte stores
    if (sizeof(TC) == 8 && vec_ok && j0 + 3 < m) {
        typedef double d2 __attribute__((ext_vector_type(2)));
        *reinterpret_cast<d2 *>(dst) = d2{(double)r[0], (double)r[1]};
        *reinterpret_cast<d2 *>(dst + 2) = d2{(double)r[2], (double)r[3]};
        return;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e)
        if (j0 + e < m) dst[e] = from_res<TC>(r[e]);
}

// One plane (q_0 = p <= 255): E holds C mod p itself.  Each wave takes 1024
// rows of one column, lanes along the rows (64 consecutive bytes per load,
// 512 bytes per 8-byte store).
template <typename TC, int EPI>
__global__ __launch_bounds__(256) void k_combine1(int64_t m, int64_t n, const uint8_t *__restrict__ E, int64_t Mp,
                                                 TC *__restrict__ C, int64_t ldc, ModP mp,
                                                 const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const int64_t chunks = (m + 1023) / 1024;
    const int64_t wv = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t i = wv / chunks;
    if (i >= n) return;
    const int64_t base = (wv % chunks) * 1024 + (threadIdx.x & 63);
    const uint8_t *src = E + i * Mp;
    TC *dst = C + i * ldc;
    if (EPI == EPI_STORE && base - (threadIdx.x & 63) + 1024 <= m) {
        // all 1024 rows inside the column: the 16 loads go out together
        uint32_t r[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) r[e] = src[base + 64 * e];
#pragma unroll
        for (int e = 0; e < 16; ++e) dst[base + 64 * e] = from_res<TC>(r[e]);
        return;
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int64_t row = base + 64 * e;
        if (row < m) {
            uint32_t r = src[row];
            if (EPI == EPI_SUB) r = submod(to_res<TC>(dst[row], mp), r, mp);
            else if (EPI == EPI_ADD) r = addmod(to_res<TC>(dst[row], mp), r, mp);
            dst[row] = from_res<TC>(r);
        }
    }
}

}  // namespace rns
