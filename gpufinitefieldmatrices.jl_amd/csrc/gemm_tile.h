// gemm_tile.h -- one BM x BN output tile of the plain FP64-MFMA mod-p GEMM,
// shared by the single-product kernel and the batched persistent kernel.
//
// Operands are reduced on load and mapped to balanced representatives in
// [-floor(p/2), floor(p/2)]; v_mfma_f64_16x16x4f64 adds four products per
// step; every R steps the accumulators are brought back into (-p, p) with
// acc - p*rint(acc/p), so |acc| <= p + R*4*floor(p/2)^2 < 2^53 and every
// partial sum is an exact integer.
#pragma once
#include "common.h"

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int PBK = 16;  // k per LDS stage

template <int BM, int BN> struct F64Smem {
    double As[PBK][BM + 4];
    double Bs[PBK][BN + 4];
};

template <typename TA, typename TB, typename TC, int BM, int BN, int EPI>
__device__ __forceinline__ void gemm_f64_tile(int64_t m, int64_t n, int64_t k, const TA *__restrict__ A, int64_t lda,
                                              const TB *__restrict__ B, int64_t ldb, TC *__restrict__ C, int64_t ldc,
                                              const ModP &mp, int R, int64_t r0, int64_t c0, F64Smem<BM, BN> &sm)
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

    double ra[APT], rb[BPT];
    const double hp = (double)mp.half;

    auto load = [&](int64_t k0) {
#pragma unroll
        for (int i = 0; i < APT; ++i) {
            int idx = tid + 256 * i;
            int row = idx % BM, kk = idx / BM;
            int64_t gr = r0 + row, gk = k0 + kk;
            double v = 0.0;
            if (gr < m && gk < k) {
                double x = (double)to_res<TA>(A[gr + gk * lda], mp);
                v = x > hp ? x - mp.pd : x;
            }
            ra[i] = v;
        }
#pragma unroll
        for (int i = 0; i < BPT; ++i) {
            int idx = tid + 256 * i;
            int kk = idx % PBK, col = idx / PBK;
            int64_t gk = k0 + kk, gc = c0 + col;
            double v = 0.0;
            if (gk < k && gc < n) {
                double x = (double)to_res<TB>(B[gk + gc * ldb], mp);
                v = x > hp ? x - mp.pd : x;
            }
            rb[i] = v;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int i = 0; i < APT; ++i) {
            int idx = tid + 256 * i;
            sm.As[idx / BM][idx % BM] = ra[i];
        }
#pragma unroll
        for (int i = 0; i < BPT; ++i) {
            int idx = tid + 256 * i;
            sm.Bs[idx % PBK][idx / PBK] = rb[i];
        }
    };
    auto reduce_all = [&]() {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    double x = acc[i][j][e];
                    acc[i][j][e] = fma(-rint(x * mp.pinv), mp.pd, x);
                }
    };

    int cnt = 0;
    const int64_t nk = (k + PBK - 1) / PBK;
    if (nk > 0) load(0);
    for (int64_t kt = 0; kt < nk; ++kt) {
        stash();
        __syncthreads();
        if (kt + 1 < nk) load((kt + 1) * PBK);
#pragma unroll
        for (int s = 0; s < PBK / 4; ++s) {
            double a[FM], b[FN];
            const int kk = s * 4 + (lane >> 4);
#pragma unroll
            for (int i = 0; i < FM; ++i) a[i] = sm.As[kk][wm * WM + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < FN; ++j) b[j] = sm.Bs[kk][wn * WN + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
            if (++cnt == R) {
                reduce_all();
                cnt = 0;
            }
        }
        __syncthreads();
    }

    // epilogue: for v_mfma_f64_16x16x4f64 the lane holds column (lane & 15),
    // rows (lane >> 4) + 4*e
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            int64_t gc = c0 + wn * WN + j * 16 + (lane & 15);
            if (gc >= n) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                int64_t gr = r0 + wm * WM + i * 16 + (lane >> 4) + 4 * e;
                if (gr >= m) continue;
                uint32_t r = (uint32_t)dmod(acc[i][j][e], mp);
                TC *dst = C + gr + gc * ldc;
                if (EPI == EPI_SUB) r = submod(to_res<TC>(*dst, mp), r, mp);
                else if (EPI == EPI_ADD) r = addmod(to_res<TC>(*dst, mp), r, mp);
                *dst = from_res<TC>(r);
            }
        }
}

// reduction interval: |acc| <= p + R*4*h^2 < 2^53 with h = floor(p/2)
static inline int f64_reduce_interval(int64_t p)
{
    double h = (double)(p / 2);
    double lim = 9007199254740992.0 - (double)p;
    double Rd = h > 0 ? floor(lim / (4.0 * h * h)) : 1e9;
    if (Rd < 1) return 0;
    return Rd > 1e9 ? 1000000000 : (int)Rd;
}
