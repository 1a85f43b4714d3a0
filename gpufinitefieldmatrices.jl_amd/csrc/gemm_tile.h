// gemm_tile.h -- one BM x BN output tile of the plain FP64-MFMA mod-p GEMM,
// shared by the single-product kernel and the batched persistent kernel.
//
// Operands are reduced on load to balanced representatives of magnitude at
// most (p+1)/2 (one rounded quotient each); v_mfma_f64_16x16x4f64 adds four
// products per step; every R steps the accumulators are brought back into
// (-p, p) with acc - p*rint(acc/p), so |acc| <= p + R*4*((p+1)/2)^2 <= 2^53
// and every partial sum is an exact integer.
#pragma once
#include "common.h"

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int PBK = 16;  // k per LDS stage

// two LDS stages: stage k+1 is stored while stage k feeds the MFMAs, one
// barrier per stage
template <int BM, int BN> struct F64Smem {
    double As[2][PBK][BM + 4];
    double Bs[2][PBK][BN + 4];
};

// Reduction cadence, fixed at compile time so that the accumulators can stay
// in the MFMA's accumulation registers between reductions:
//   RED_NONE   the whole contraction fits the exact range (k4-steps <= R)
//   RED_STEP   every k4-step (R = 1)
//   RED_2      every second k4-step (R = 2 or 3)
//   RED_STAGE  after every rq-th 16-deep stage (rq = R / 4, R >= 4)
enum { RED_NONE = 0, RED_STEP = 1, RED_2 = 2, RED_STAGE = 4 };

static inline int f64_reduce_mode(int R, int64_t k, int *rq)
{
    *rq = R >= 4 ? R / 4 : 1;
    if ((int64_t)R >= (k + 3) / 4) return RED_NONE;
    if (R >= 4) return RED_STAGE;
    if (R >= 2) return RED_2;
    return RED_STEP;
}

// balanced representative of a stored element: for the floating types one
// rounded quotient, x - p rint(x/p), which for integral |x| < 2^51 lies in
// [-(p+1)/2, (p+1)/2] (the rounded quotient is within 1/2 + 1/(2p) of x/p);
// int32 working arrays already hold residues in [0, p)
template <typename T> __device__ __forceinline__ double bal_of(T x, const ModP &mp)
{
    const double d = (double)x;
    return fma(-rint(d * mp.pinv), mp.pd, d);
}
template <> __device__ __forceinline__ double bal_of<int32_t>(int32_t x, const ModP &mp)
{
    const double d = (double)x;
    return d > (double)mp.half ? d - mp.pd : d;
}

template <typename TA, typename TB, typename TC, int BM, int BN, int EPI, int RED>
__device__ __forceinline__ void gemm_f64_tile(int64_t m, int64_t n, int64_t k, const TA *__restrict__ A, int64_t lda,
                                              const TB *__restrict__ B, int64_t ldb, TC *__restrict__ C, int64_t ldc,
                                              const ModP &mp, int rq, int64_t r0, int64_t c0, F64Smem<BM, BN> &sm)
{
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int APT = BM * PBK / 256;  // A elements per thread per stage
    constexpr int BPT = BN * PBK / 256;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    v4d acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    // a stage's raw elements, loaded at indices clamped into the operands
    // (so no load waits behind a branch) with a bit per element that is in
    // range; they are reduced and masked when stashed, one stage later
    static_assert(APT <= 16 && BPT <= 16, "validity mask");
    TA xa[APT];
    TB xb[BPT];
    uint32_t vmask = 0;

    auto load = [&](int64_t k0) {
        vmask = 0;
#pragma unroll
        for (int i = 0; i < APT; ++i) {
            int idx = tid + 256 * i;
            int row = idx % BM, kk = idx / BM;
            int64_t gr = r0 + row, gk = k0 + kk;
            xa[i] = A[min(gr, m - 1) + min(gk, k - 1) * lda];
            vmask |= (gr < m && gk < k) ? 1u << i : 0u;
        }
#pragma unroll
        for (int i = 0; i < BPT; ++i) {
            int idx = tid + 256 * i;
            int kk = idx % PBK, col = idx / PBK;
            int64_t gk = k0 + kk, gc = c0 + col;
            xb[i] = B[min(gk, k - 1) + min(gc, n - 1) * ldb];
            vmask |= (gk < k && gc < n) ? 1u << (16 + i) : 0u;
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int i = 0; i < APT; ++i) {
            int idx = tid + 256 * i;
            sm.As[buf][idx / BM][idx % BM] = ((vmask >> i) & 1u) ? bal_of<TA>(xa[i], mp) : 0.0;
        }
#pragma unroll
        for (int i = 0; i < BPT; ++i) {
            int idx = tid + 256 * i;
            sm.Bs[buf][idx % PBK][idx / PBK] = ((vmask >> (16 + i)) & 1u) ? bal_of<TB>(xb[i], mp) : 0.0;
        }
    };

    auto reduce_all = [&]() {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const double x = acc[i][j][e];
                    acc[i][j][e] = fma(-rint(x * mp.pinv), mp.pd, x);
                }
    };

    const int64_t nk = (k + PBK - 1) / PBK;
    if (nk > 0) {
        load(0);
        stash(0);
        if (nk > 1) load(PBK);
    }
    __syncthreads();
    int sc = 0;  // stages since the last reduction (RED_STAGE)
    for (int64_t kt = 0; kt < nk; ++kt) {
        const int cur = (int)(kt & 1);
        // buffer cur ^ 1 was last read before the previous barrier
        if (kt + 1 < nk) stash(cur ^ 1);
        if (kt + 2 < nk) load((kt + 2) * PBK);
#pragma unroll
        for (int s = 0; s < PBK / 4; ++s) {
            double a[FM], b[FN];
            const int kk = s * 4 + (lane >> 4);
#pragma unroll
            for (int i = 0; i < FM; ++i) a[i] = sm.As[cur][kk][wm * WM + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < FN; ++j) b[j] = sm.Bs[cur][kk][wn * WN + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
            if (RED == RED_STEP || (RED == RED_2 && (s & 1))) reduce_all();
        }
        if (RED == RED_STAGE && ++sc == rq) {
            sc = 0;
            reduce_all();
        }
        __syncthreads();
    }

    // epilogue through LDS: for v_mfma_f64_16x16x4f64 the lane holds column
    // (lane & 15), rows (lane >> 4) + 4*e, so direct stores would touch 16
    // columns per instruction.  The reduced residues are transposed into a
    // column-major LDS tile (the operand stages are free after the last
    // barrier), and each wave then reads and writes 64 consecutive rows of C.
    constexpr int TP = BM + 1;  // LDS column pitch (odd: no bank conflicts)
    static_assert(sizeof(F64Smem<BM, BN>) >= sizeof(uint32_t) * BN * TP, "LDS tile too small");
    uint32_t *T = reinterpret_cast<uint32_t *>(&sm);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int cl = wn * WN + j * 16 + (lane & 15);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int rl = wm * WM + i * 16 + (lane >> 4) + 4 * e;
                T[cl * TP + rl] = (uint32_t)dmod(acc[i][j][e], mp);
            }
        }
    __syncthreads();
    constexpr int EPT = BM * BN / 256;
#pragma unroll
    for (int x = 0; x < EPT; ++x) {
        const int idx = tid + 256 * x;
        const int rl = idx % BM, cl = idx / BM;
        const int64_t gr = r0 + rl, gc = c0 + cl;
        if (gr < m && gc < n) {
            uint32_t r = T[cl * TP + rl];
            TC *dst = C + gr + gc * ldc;
            if (EPI == EPI_SUB) r = submod(to_res<TC>(*dst, mp), r, mp);
            else if (EPI == EPI_ADD) r = addmod(to_res<TC>(*dst, mp), r, mp);
            *dst = from_res<TC>(r);
        }
    }
    // the LDS tile is reused by the caller's next tile
    __syncthreads();
}

// reduction interval: |acc| <= p + R*4*h^2 < 2^53 with h = floor(p/2)
static inline int f64_reduce_interval(int64_t p)
{
    // operands are balanced to at most (p+1)/2 in magnitude (bal_of)
    double h = (double)((p + 1) / 2);
    double lim = 9007199254740992.0 - (double)p;
    double Rd = h > 0 ? floor(lim / (4.0 * h * h)) : 1e9;
    if (Rd < 1) return 0;
    return Rd > 1e9 ? 1000000000 : (int)Rd;
}
