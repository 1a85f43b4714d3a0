// batched.hip -- thousands of independent small products / eliminations, each
// under its own prime, in one persistent launch fed by a device work queue.
//
// GEMM: the plan sorts jobs by cost, cuts each into output tiles whose shape
// the autotuner picks for the job's shape bucket, and captures
// [queue reset kernel, persistent kernel] in a hipGraph; every workgroup pulls
// tile indices from an atomic counter until the queue is empty.
//
// Elimination: one wavefront per instance (m, n <= 128) with the matrix held
// in LDS, the same pivot rule as the single-matrix path; a plan uploads the
// job table once and captures the persistent kernel in a hipGraph.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <numeric>
#include <type_traits>

#include "batched.h"
#include "gemm_limb.h"
#include "gemm_tile.h"

// ---------------------------------------------------------------------------
// batched GEMM
// ---------------------------------------------------------------------------

struct BJob {
    const void *A; int64_t lda;
    const void *B; int64_t ldb;
    void *C; int64_t ldc;
    int m, n, k, R;
    int tile;     // 32 or 64
    int tiles_n;  // tiles along n
    ModP mp;
    int L;        // digit form: digits per residue (0: FP64 form)
    int lazy;     // digit form: the recombination needs no reduction per term
    double w[7];  // digit form: 256^s mod p
};

struct BItem {
    int job;
    int tile;  // tile index within the job: ti = tile / tiles_n, tj = tile % tiles_n
};

// Work items are grouped in classes: FP64 form (tile 64 or 32) x (reduction
// needed or not), and the int8 digit form by digit count (1..4); each class
// is served by its own persistent kernel, which carries only that class's
// tile code (fewer registers, more resident waves).
// Within a class every product is given to one of the 8 XCDs (balanced by
// cost) and its tiles go to that XCD's queue: the workgroups of an XCD (the
// dispatcher deals workgroups to the XCDs in turn, so workgroup w runs on
// XCD w mod 8) take their tiles first, so the tiles of one product share
// its operands in one L2; a workgroup whose queue is empty then helps the
// others.
// 0-3 FP64 form, 4-7 digit form with 1-4 digits (a 64 x 64 tile per item),
// 8-11 digit form with 1-4 digits for k <= 256 (a 64-row panel per item)
constexpr int NCLASS = 12;
constexpr int NXCD = 8;
struct XcdRanges {
    int64_t off[NXCD + 1];  // item ranges of the XCD queues within a class
};
struct modp_gemm_plan {
    int kind = EK_F64;
    int64_t count = 0, items = 0;
    int64_t cls_off[NCLASS + 1] = {0};  // item ranges of the classes
    XcdRanges xr[NCLASS];               // XCD queues of each class (relative to cls_off)
    BJob *djobs = nullptr;
    BItem *ditems = nullptr;
    int *dcounter = nullptr;  // one counter per class and XCD, 64 bytes apart
    int grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

__global__ void k_queue_reset(int *counter)
{
    if (threadIdx.x < NCLASS * NXCD && blockIdx.x == 0) counter[threadIdx.x * 16] = 0;
}

template <typename T, int TILE, int RED>
__global__ __launch_bounds__(256, 2) void k_batched_gemm(const BJob *__restrict__ jobs,
                                                         const BItem *__restrict__ items, XcdRanges xr,
                                                         int *__restrict__ counter)
{
    __shared__ F64Smem<TILE, TILE> sm;
    __shared__ int s_item;
    const int home = blockIdx.x % NXCD;
    for (int qq = 0; qq < NXCD; ++qq) {
        const int q = (home + qq) % NXCD;
        const int64_t base = xr.off[q], cnt = xr.off[q + 1] - base;
        if (cnt <= 0) continue;
        for (;;) {
            if (threadIdx.x == 0) s_item = atomicAdd(counter + 16 * q, 1);
            __syncthreads();
            const int it = s_item;
            __syncthreads();
            if (it >= cnt) break;
            const BItem item = items[base + it];
            const BJob &j = jobs[item.job];
            const int ti = item.tile / j.tiles_n, tj = item.tile % j.tiles_n;
            gemm_f64_tile<T, T, T, TILE, TILE, EPI_STORE, RED>(j.m, j.n, j.k, (const T *)j.A, j.lda,
                                                               (const T *)j.B, j.ldb, (T *)j.C, j.ldc, j.mp, 1,
                                                               (int64_t)ti * TILE, (int64_t)tj * TILE, sm);
        }
    }
}

// Digit form of one work item: a 64 x 64 output tile on four waves (32 x 32
// blocks, v_mfma_i32_32x32x32_i8).  Per 32-deep stage every thread loads 8
// elements of A (one row, the lanes along the rows: 512-byte runs) and 8 of B
// (lane pairs of consecutive k, the lanes along k: four 256-byte column runs
// per load), turns each residue into L signed base-256 digits (the limb
// form's representation) and writes them to the digit planes in LDS; the
// next stage's loads are in flight during the MFMAs.  The L x L digit
// products accumulate by digit sum s = u + v in 2L - 1 int32 accumulators
// (|sum| <= L k 2^14 < 2^31 for k <= 16384), and the epilogue forms
// sum_s (acc_s mod p) (256^s mod p) mod p in FP64.  The blocks are computed
// transposed (B digits as the first operand), so lane l holds row l % 32 of
// C and a store writes 32 consecutive rows of a column.  LDS rows hold 32
// digits (2 chunks); chunk c of row r sits at slot c ^ ((r >> 3) & 1), which
// keeps the 16-lane fragment reads free of bank conflicts.
constexpr int BL_TILE = 64, BL_BK = 32, BL_KMAX = 16384;
typedef int bl_v2i __attribute__((ext_vector_type(2)));
typedef int bl_v16i __attribute__((ext_vector_type(16)));

// Residue of a stored element in the digit representation (top as in
// digit_rep): a compare when it is already in [0, p) (the usual case), the
// FP64 reduction otherwise.
template <typename T>
__device__ __forceinline__ int32_t bl_residue(T v, const ModP &mp, int32_t top)
{
    const double x = (double)v;
    uint32_t r = (uint32_t)x;
    if (!(x >= 0.0 && x < mp.pd)) r = (uint32_t)dmod(x, mp);
    return digit_rep(r, top, mp);
}

// Digits: the stored byte of digit u of x is the low byte of x_u, with
// x_0 = x and x_{u+1} = (x_u + 128) >> 8 (arithmetic), the signed base-256
// digits of the limb form.  Eight consecutive k of one A row: digit u goes to
// dst(u) as one 8-byte word.
template <int L, class Dst> __device__ __forceinline__ void bl_store_a8(int32_t (&x)[8], Dst dst)
{
#pragma unroll
    for (int u = 0; u < L; ++u) {
        bl_v2i w;
#pragma unroll
        for (int d = 0; d < 2; ++d)
            w[d] = (int)(((uint32_t)x[4 * d] & 255u) | (((uint32_t)x[4 * d + 1] & 255u) << 8) |
                         (((uint32_t)x[4 * d + 2] & 255u) << 16) | ((uint32_t)x[4 * d + 3] << 24));
        *reinterpret_cast<bl_v2i *>(dst(u)) = w;
        if (u + 1 < L) {
#pragma unroll
            for (int e = 0; e < 8; ++e) x[e] = (x[e] + 128) >> 8;
        }
    }
}

// B stage of 32 k into the digit planes Bs[u] (64 columns of 32 bytes, chunk
// c of column r at slot c ^ ((r >> 3) & 1)): element 2 i + t of vb is k pair
// bk2 + t of column 16 wave + 4 i + bq.
template <typename T, int L>
__device__ __forceinline__ void bl_store_b(const T (&vb)[8], int8_t (*Bs)[BL_TILE * BL_BK], int wave, int bq,
                                           int bk2, const ModP &mp, int32_t top)
{
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int c = 16 * wave + 4 * i + bq;
        const int off = c * BL_BK + 16 * ((bk2 >> 4) ^ ((c >> 3) & 1)) + (bk2 & 15);
        int32_t x0 = bl_residue(vb[2 * i], mp, top), x1 = bl_residue(vb[2 * i + 1], mp, top);
#pragma unroll
        for (int u = 0; u < L; ++u) {
            *reinterpret_cast<uint16_t *>(&Bs[u][off]) = (uint16_t)(((uint32_t)x0 & 255u) | (((uint32_t)x1 & 255u) << 8));
            x0 = (x0 + 128) >> 8;
            x1 = (x1 + 128) >> 8;
        }
    }
}

// Epilogue of one 32 x 32 block held transposed (lane l: row i = l % 32 of C,
// columns j0 + 8 (e / 4) + 4 (l / 32) + e % 4 of element e):
// sum_s (acc_s mod p) (256^s mod p) mod p in FP64, stored as T.
template <typename T, int NA>
__device__ __forceinline__ void bl_store_block(const bl_v16i (&acc)[NA], const BJob &j, int64_t i, int64_t j0, int h)
{
    const ModP mp = j.mp;
    T *C = (T *)j.C;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int64_t jj = j0 + 8 * (e >> 2) + 4 * h + (e & 3);
        double S = 0.0;
        if (j.lazy) {
            // |S| < 2^51: the rounded quotient is within one of the exact
            // one, so one fix-up gives S mod p
#pragma unroll
            for (int t = 0; t < NA; ++t) S = fma((double)acc[t][e], j.w[t], S);
            S = fma(-rint(S * mp.pinv), mp.pd, S);
            S = S < 0.0 ? S + mp.pd : S;
        } else {
#pragma unroll
            for (int t = 0; t < NA; ++t) S = dmod(fma(dmod((double)acc[t][e], mp), j.w[t], S), mp);
        }
        // S is an exact integer in [0, p)
        if (i < j.m && jj < j.n) C[i + jj * j.ldc] = (T)S;
    }
}

template <typename T, int L>
__global__ __launch_bounds__(256, 2) void k_batched_limb(const BJob *__restrict__ jobs,
                                                         const BItem *__restrict__ items, XcdRanges xr,
                                                         int *__restrict__ counter)
{
    typedef int v4i __attribute__((ext_vector_type(4)));
    typedef bl_v16i v16i;
    constexpr int TL = BL_TILE, BK = BL_BK, NA = 2 * L - 1;
    constexpr int32_t TOP = (int32_t)(127 * ((((int64_t)1) << (8 * L)) - 1) / 255);
    __shared__ __attribute__((aligned(16))) int8_t As[2][L][TL * BK];
    __shared__ __attribute__((aligned(16))) int8_t Bs[2][L][TL * BK];
    __shared__ int s_item;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave & 1, wc = wave >> 1;
    // A staging: row sr, k quarter sc (8 k); B staging: column 16 wave +
    // 4 i + lane / 16 (i < 4), k pair 2 (lane % 16)
    const int sr = tid & 63, sc = tid >> 6;
    const int aslot = sr * BK + 16 * ((sc >> 1) ^ ((sr >> 3) & 1)) + 8 * (sc & 1);
    const int bk2 = 2 * (lane & 15), bq = lane >> 4;
    const int fr = lane & 31, h = lane >> 5, fsw = (fr >> 3) & 1;
    const int home = blockIdx.x % NXCD;
    for (int qq = 0; qq < NXCD; ++qq) {
        const int q = (home + qq) % NXCD;
        const int64_t base = xr.off[q], cnt = xr.off[q + 1] - base;
        if (cnt <= 0) continue;
        for (;;) {
            if (threadIdx.x == 0) s_item = atomicAdd(counter + 16 * q, 1);
            __syncthreads();
            const int it = s_item;
            __syncthreads();
            if (it >= cnt) break;
            const BItem item = items[base + it];
            const BJob &j = jobs[item.job];
            const int ti = item.tile / j.tiles_n, tj = item.tile % j.tiles_n;
            const int64_t r0 = (int64_t)ti * TL, c0 = (int64_t)tj * TL;
            const T *A = (const T *)j.A, *B = (const T *)j.B;
            const ModP mp = j.mp;
            const int K = j.k, ns = (K + BK - 1) / BK;
            const int64_t arow = r0 + sr;
            const bool arow_ok = arow < j.m;
            T va[8], vb[8];
            auto gload = [&](int st) {
                const int64_t kb = (int64_t)st * BK;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int64_t kk = kb + 8 * sc + e;
                    va[e] = (arow_ok && kk < K) ? A[arow + kk * j.lda] : (T)0;
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int64_t col = c0 + 16 * wave + 4 * i + bq;
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        const int64_t kk = kb + bk2 + t;
                        vb[2 * i + t] = (col < j.n && kk < K) ? B[kk + col * j.ldb] : (T)0;
                    }
                }
            };
            auto lstore = [&](int buf) {
                int32_t xa[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) xa[e] = bl_residue(va[e], mp, TOP);
                bl_store_a8<L>(xa, [&](int u) { return &As[buf][u][aslot]; });
                bl_store_b<T, L>(vb, Bs[buf], wave, bq, bk2, mp, TOP);
            };
            v16i acc[NA];
#pragma unroll
            for (int t = 0; t < NA; ++t) acc[t] = v16i{0};
            gload(0);
            lstore(0);
            __syncthreads();
            for (int st = 0; st < ns; ++st) {
                const int buf = st & 1;
                if (st + 1 < ns) gload(st + 1);
#pragma unroll
                for (int kk = 0; kk < BK / 32; ++kk) {
                    const int off = 16 * ((2 * kk + h) ^ fsw);  // BK = 32: kk = 0, chunk h
                    v4i fa[L], fb[L];
#pragma unroll
                    for (int u = 0; u < L; ++u) {
                        fa[u] = *reinterpret_cast<const v4i *>(&As[buf][u][(wr * 32 + fr) * BK + off]);
                        fb[u] = *reinterpret_cast<const v4i *>(&Bs[buf][u][(wc * 32 + fr) * BK + off]);
                    }
#pragma unroll
                    for (int u = 0; u < L; ++u)
#pragma unroll
                        for (int v = 0; v < L; ++v)
                            acc[u + v] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fb[v], fa[u], acc[u + v], 0, 0, 0);
                }
                if (st + 1 < ns) lstore(buf ^ 1);
                __syncthreads();
            }
            bl_store_block<T, NA>(acc, j, r0 + wr * 32 + fr, c0 + wc * 32, h);
        }
    }
}

// Digit form, one 64-row panel per work item (k <= 256): the panel's rows of
// A are split into digits during its first column tile, stage by stage in the
// same software pipeline as B, and kept in LDS for the later column tiles
// (rows of 256 + 16 bytes: 16 consecutive rows read at one offset fall in
// distinct bank groups).  The A digits of a product are computed once per row
// panel instead of once per tile.
constexpr int RP_KMAX = 256;
template <typename T, int L>
__global__ __launch_bounds__(256, L == 4 ? 1 : 2) void k_batched_limb_rp(const BJob *__restrict__ jobs,
                                                            const BItem *__restrict__ items, XcdRanges xr,
                                                            int *__restrict__ counter)
{
    typedef int v4i __attribute__((ext_vector_type(4)));
    typedef bl_v16i v16i;
    constexpr int TL = BL_TILE, BK = BL_BK, NA = 2 * L - 1, AP = RP_KMAX + 16;
    constexpr int32_t TOP = (int32_t)(127 * ((((int64_t)1) << (8 * L)) - 1) / 255);
    __shared__ __attribute__((aligned(16))) int8_t As[L][TL * AP];
    __shared__ __attribute__((aligned(16))) int8_t Bs[2][L][TL * BK];
    __shared__ int s_item;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave & 1, wc = wave >> 1;
    const int sr = tid & 63, sc = tid >> 6;
    const int bk2 = 2 * (lane & 15), bq = lane >> 4;
    const int fr = lane & 31, h = lane >> 5, fsw = (fr >> 3) & 1;
    const int home = blockIdx.x % NXCD;
    for (int qq = 0; qq < NXCD; ++qq) {
        const int q = (home + qq) % NXCD;
        const int64_t base = xr.off[q], cnt = xr.off[q + 1] - base;
        if (cnt <= 0) continue;
        for (;;) {
            if (threadIdx.x == 0) s_item = atomicAdd(counter + 16 * q, 1);
            __syncthreads();
            const int it = s_item;
            __syncthreads();
            if (it >= cnt) break;
            const BItem item = items[base + it];
            const BJob &j = jobs[item.job];
            const int64_t r0 = (int64_t)item.tile * TL;
            const T *A = (const T *)j.A, *B = (const T *)j.B;
            const ModP mp = j.mp;
            const int K = j.k, ns = (K + BK - 1) / BK;
            const int64_t arow = r0 + sr;
            const bool arow_ok = arow < j.m;
            for (int tj = 0; tj < j.tiles_n; ++tj) {
                const int64_t c0 = (int64_t)tj * TL;
                // the first column tile also stages the A panel: thread (row
                // sr, quarter sc) splits k = 32 st + 8 sc .. + 7 of each stage
                const bool first = tj == 0;
                T va[8], vb[8];
                auto gload = [&](int st) {
                    const int64_t kb = (int64_t)st * BK;
                    if (first) {
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const int64_t kk = kb + 8 * sc + e;
                            va[e] = (arow_ok && kk < K) ? A[arow + kk * j.lda] : (T)0;
                        }
                    }
#pragma unroll
                    for (int ii = 0; ii < 4; ++ii) {
                        const int64_t col = c0 + 16 * wave + 4 * ii + bq;
#pragma unroll
                        for (int t = 0; t < 2; ++t) {
                            const int64_t kk = kb + bk2 + t;
                            vb[2 * ii + t] = (col < j.n && kk < K) ? B[kk + col * j.ldb] : (T)0;
                        }
                    }
                };
                auto lstore = [&](int buf, int st) {
                    if (first) {
                        int32_t xa[8];
#pragma unroll
                        for (int e = 0; e < 8; ++e) xa[e] = bl_residue(va[e], mp, TOP);
                        bl_store_a8<L>(xa, [&](int u) { return &As[u][sr * AP + st * BK + 8 * sc]; });
                    }
                    bl_store_b<T, L>(vb, Bs[buf], wave, bq, bk2, mp, TOP);
                };
                v16i acc[NA];
#pragma unroll
                for (int t = 0; t < NA; ++t) acc[t] = v16i{0};
                if (ns > 0) {
                    gload(0);
                    lstore(0, 0);
                }
                __syncthreads();
                for (int st = 0; st < ns; ++st) {
                    const int buf = st & 1;
                    if (st + 1 < ns) gload(st + 1);
                    v4i fa[L], fb[L];
#pragma unroll
                    for (int u = 0; u < L; ++u) {
                        fa[u] = *reinterpret_cast<const v4i *>(&As[u][(wr * 32 + fr) * AP + st * BK + 16 * h]);
                        fb[u] = *reinterpret_cast<const v4i *>(&Bs[buf][u][(wc * 32 + fr) * BK + 16 * (h ^ fsw)]);
                    }
#pragma unroll
                    for (int u = 0; u < L; ++u)
#pragma unroll
                        for (int v = 0; v < L; ++v)
                            acc[u + v] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fb[v], fa[u], acc[u + v], 0, 0, 0);
                    if (st + 1 < ns) lstore(buf ^ 1, st + 1);
                    __syncthreads();
                }
                bl_store_block<T, NA>(acc, j, r0 + wr * 32 + fr, c0 + wc * 32, h);
            }
        }
    }
}

template <typename T>
void launch_class(int cls, const modp_gemm_plan *pl, hipStream_t st)
{
    const int64_t n = pl->cls_off[cls + 1] - pl->cls_off[cls];
    if (n <= 0) return;
    const BItem *it = pl->ditems + pl->cls_off[cls];
    int *ctr = pl->dcounter + cls * 16 * NXCD;
    const XcdRanges &xr = pl->xr[cls];
    // resident workgroups per CU by the kernels' register and LDS use: 3 for
    // 64x64 FP64 tiles, 5 for 32x32; 4, 3, 2, 2 for the digit form with 1-4
    // digits; 4, 3, 2, 1 for its row-panel kernel
    static const int per_cu[NCLASS] = {3, 3, 5, 5, 4, 3, 2, 2, 4, 3, 2, 1};
    const int64_t cap = (int64_t)pl->grid * per_cu[cls];
    // a multiple of the XCD count, so every XCD gets the same number
    int64_t g64 = n < cap ? n : cap;
    g64 = (g64 + NXCD - 1) / NXCD * NXCD;
    const unsigned g = (unsigned)g64;
    switch (cls) {
    case 0: k_batched_gemm<T, 64, RED_NONE><<<g, 256, 0, st>>>(pl->djobs, it, xr, ctr); break;
    case 1: k_batched_gemm<T, 64, RED_STEP><<<g, 256, 0, st>>>(pl->djobs, it, xr, ctr); break;
    case 2: k_batched_gemm<T, 32, RED_NONE><<<g, 256, 0, st>>>(pl->djobs, it, xr, ctr); break;
    case 3: k_batched_gemm<T, 32, RED_STEP><<<g, 256, 0, st>>>(pl->djobs, it, xr, ctr); break;
    case 4: k_batched_limb<T, 1><<<g, 256, 0, st>>>(pl->djobs, it, xr, ctr); break;
    case 5: k_batched_limb<T, 2><<<g, 256, 0, st>>>(pl->djobs, it, xr, ctr); break;
    case 6: k_batched_limb<T, 3><<<g, 256, 0, st>>>(pl->djobs, it, xr, ctr); break;
    case 7: k_batched_limb<T, 4><<<g, 256, 0, st>>>(pl->djobs, it, xr, ctr); break;
    case 8: k_batched_limb_rp<T, 1><<<g, 256, 0, st>>>(pl->djobs, it, xr, ctr); break;
    case 9: k_batched_limb_rp<T, 2><<<g, 256, 0, st>>>(pl->djobs, it, xr, ctr); break;
    case 10: k_batched_limb_rp<T, 3><<<g, 256, 0, st>>>(pl->djobs, it, xr, ctr); break;
    default: k_batched_limb_rp<T, 4><<<g, 256, 0, st>>>(pl->djobs, it, xr, ctr); break;
    }
}

// Tile choice per job: a 64x64 tile does 4x the MFMA work of a 32x32 tile per
// queue slot and reuses each loaded operand twice as often; take it while its
// padded work stays within thr times that of 32x32 tiles.  thr is the
// autotuned knob of a plan (0: all 32x32, large: all 64x64).
int pick_tile(int m, int n, double thr)
{
    const int64_t w64 = (int64_t)((m + 63) / 64) * ((n + 63) / 64) * 64 * 64;
    const int64_t w32 = (int64_t)((m + 31) / 32) * ((n + 31) / 32) * 32 * 32;
    return (double)w64 <= thr * (double)w32 ? 64 : 32;
}

int ceil_log2(int64_t x)
{
    int b = 0;
    while (((int64_t)1 << b) < x && b < 62) ++b;
    return b;
}

int plan_build(modp_ctx *ctx, int kind, const modp_gemm_job *jobs, int64_t count, double thr, modp_gemm_plan **out);

double plan_time(modp_ctx *ctx, modp_gemm_plan *pl, int *st)
{
    hipEvent_t a = nullptr, b = nullptr;
    double best = 1e30;
    *st = hip_check(ctx, hipEventCreate(&a), "hipEventCreate");
    if (*st == MODP_OK) *st = hip_check(ctx, hipEventCreate(&b), "hipEventCreate");
    if (*st == MODP_OK) *st = batched_gemm_run(ctx, pl);
    for (int r = 0; r < 3 && *st == MODP_OK; ++r) {
        *st = hip_check(ctx, hipEventRecord(a, ctx->stream), "hipEventRecord");
        if (*st == MODP_OK) *st = batched_gemm_run(ctx, pl);
        if (*st == MODP_OK) *st = hip_check(ctx, hipEventRecord(b, ctx->stream), "hipEventRecord");
        if (*st == MODP_OK) *st = hip_check(ctx, hipEventSynchronize(b), "hipEventSynchronize");
        float t = 0.f;
        if (*st == MODP_OK) *st = hip_check(ctx, hipEventElapsedTime(&t, a, b), "hipEventElapsedTime");
        if (*st == MODP_OK) best = std::min(best, (double)t);
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    return best;
}

}  // namespace

// Batch signature for the tuning cache: kind + 16, log2 buckets of the count
// and of the mean m, n, k.
TuneKey tune_key_batch(int kind, const modp_gemm_job *jobs, int64_t count)
{
    double sm = 0, sn = 0, sk = 0;
    for (int64_t i = 0; i < count; ++i) { sm += jobs[i].m; sn += jobs[i].n; sk += jobs[i].k; }
    const double c = count > 0 ? (double)count : 1.0;
    TuneKey k;
    k.v[0] = 16 + kind;
    k.v[1] = ceil_log2(count);
    k.v[2] = ceil_log2((int64_t)(sm / c + 0.5));
    k.v[3] = ceil_log2((int64_t)(sn / c + 0.5));
    k.v[4] = ceil_log2((int64_t)(sk / c + 0.5));
    k.v[5] = 0;
    k.v[6] = 0;
    return k;
}

// The plan's tile policy comes from the tuning cache; on a miss with
// MODP_TUNE_ON_MISS the candidate thresholds are built, each run (every run
// rewrites the same outputs) and timed, and the fastest plan is kept.
int batched_gemm_plan(modp_ctx *ctx, int kind, const modp_gemm_job *jobs, int64_t count, modp_gemm_plan **out)
{
    if (kind == EK_I32) return set_error(ctx, MODP_ERR_DTYPE, "batched gemm: unsupported element type");
    const double default_thr = 1.25;
    if (ctx->tune_mode == MODP_TUNE_OFF || count == 0) return plan_build(ctx, kind, jobs, count, default_thr, out);
    const TuneKey key = tune_key_batch(kind, jobs, count);
    auto it = ctx->tune_cache.find(key);
    if (it != ctx->tune_cache.end()) return plan_build(ctx, kind, jobs, count, it->second.thr, out);
    if (ctx->tune_mode != MODP_TUNE_ON_MISS) return plan_build(ctx, kind, jobs, count, default_thr, out);
    const double thrs[5] = {0.0, 1.0, 1.25, 1.6, 1e9};
    modp_gemm_plan *best = nullptr;
    double best_ms = 1e30, best_thr = default_thr;
    for (double thr : thrs) {
        modp_gemm_plan *pl = nullptr;
        int st = plan_build(ctx, kind, jobs, count, thr, &pl);
        if (st != MODP_OK) {
            batched_gemm_destroy(ctx, best);
            return st;
        }
        const double ms = plan_time(ctx, pl, &st);
        if (st != MODP_OK) {
            batched_gemm_destroy(ctx, pl);
            batched_gemm_destroy(ctx, best);
            return st;
        }
        if (ms < best_ms) {
            batched_gemm_destroy(ctx, best);
            best = pl;
            best_ms = ms;
            best_thr = thr;
        } else {
            batched_gemm_destroy(ctx, pl);
        }
    }
    TuneEntry e;
    e.thr = best_thr;
    e.ms = best_ms;
    ctx->tune_cache[key] = e;
    *out = best;
    return MODP_OK;
}

namespace {

int plan_build(modp_ctx *ctx, int kind, const modp_gemm_job *jobs, int64_t count, double thr, modp_gemm_plan **out)
{
    // MODP_BATCHED_FP64=1 keeps every product in the FP64 form (used to
    // compare the two forms)
    const char *fenv = getenv("MODP_BATCHED_FP64");
    const bool fp64_only = fenv && atoi(fenv) != 0;
    // MODP_BATCHED_TILES=1 keeps digit-form products on 64 x 64 tile items
    // (no row panels)
    const char *tenv = getenv("MODP_BATCHED_TILES");
    const bool tile_items = tenv && atoi(tenv) != 0;
    modp_gemm_plan *pl = new modp_gemm_plan();
    pl->kind = kind;
    pl->count = count;
    std::vector<BJob> hj((size_t)count);
    std::vector<int64_t> order((size_t)count);
    std::iota(order.begin(), order.end(), 0);
    for (int64_t i = 0; i < count; ++i) {
        const modp_gemm_job &s = jobs[i];
        BJob &d = hj[i];
        d.A = s.A; d.lda = s.lda; d.B = s.B; d.ldb = s.ldb; d.C = s.C; d.ldc = s.ldc;
        d.m = (int)s.m; d.n = (int)s.n; d.k = (int)s.k;
        d.R = f64_reduce_interval(s.p);
        d.mp = make_modp(s.p);
        d.tile = pick_tile(d.m, d.n, thr);
        d.L = 0;
        d.lazy = 0;
        const int L = limb_count(s.p);
        if (!fp64_only && d.k <= BL_KMAX && L <= 4) {
            // digit form: the recombination reduces once at the end when its
            // terms |acc_s| (256^s mod p), |acc_s| <= L k 2^14, sum below 2^51
            d.L = L;
            d.tile = BL_TILE;
            double w = 1.0;
            for (int t = 0; t < 7; ++t) {
                d.w[t] = w;
                w = std::fmod(w * 256.0, (double)s.p);
            }
            const double bound = (2.0 * L - 1) * L * (double)d.k * 16384.0 * (double)(s.p - 1);
            d.lazy = bound < 2251799813685248.0 ? 1 : 0;  // 2^51
        }
        d.tiles_n = (d.n + d.tile - 1) / d.tile;
    }
    // most expensive tiles first, within each class
    auto cost = [&](int64_t i) { return (int64_t)hj[i].tile * hj[i].tile * hj[i].k; };
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return cost(a) > cost(b); });
    auto cls_of = [&](const BJob &d) {
        if (d.L > 0) return (d.k <= RP_KMAX && !tile_items ? 7 : 3) + d.L;
        const bool exact = (int64_t)d.R >= ((int64_t)d.k + 3) / 4;
        return (d.tile == 64 ? 0 : 2) + (exact ? 0 : 1);
    };
    std::vector<BItem> items;
    for (int c = 0; c < NCLASS; ++c) {
        pl->cls_off[c] = (int64_t)items.size();
        // products to XCDs, largest first to the least loaded XCD
        std::vector<int64_t> prods;
        for (int64_t oi : order) {
            const BJob &d = hj[oi];
            if (d.m > 0 && d.n > 0 && cls_of(d) == c) prods.push_back(oi);
        }
        // items: output tiles, or row panels for the row-panel classes
        auto rows_of = [&](const BJob &d) { return (d.m + d.tile - 1) / d.tile; };
        auto tiles_of = [&](const BJob &d) { return c >= 8 ? rows_of(d) : rows_of(d) * d.tiles_n; };
        std::vector<int64_t> load(NXCD, 0);
        std::vector<std::vector<int64_t>> per(NXCD);
        for (int64_t oi : prods) {
            const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            load[x] += cost(oi) * rows_of(hj[oi]) * hj[oi].tiles_n;
            per[x].push_back(oi);
        }
        for (int x = 0; x < NXCD; ++x) {
            pl->xr[c].off[x] = (int64_t)items.size() - pl->cls_off[c];
            for (int64_t oi : per[x]) {
                const int tiles = tiles_of(hj[oi]);
                for (int t = 0; t < tiles; ++t) items.push_back(BItem{(int)oi, t});
            }
        }
        pl->xr[c].off[NXCD] = (int64_t)items.size() - pl->cls_off[c];
    }
    pl->cls_off[NCLASS] = (int64_t)items.size();
    pl->items = (int64_t)items.size();
    hipError_t e = hipMalloc(&pl->djobs, sizeof(BJob) * (count > 0 ? count : 1));
    if (e == hipSuccess) e = hipMalloc(&pl->ditems, sizeof(BItem) * (items.empty() ? 1 : items.size()));
    if (e == hipSuccess) e = hipMalloc(&pl->dcounter, sizeof(int) * 16 * NCLASS * NXCD);
    if (e == hipSuccess && count > 0)
        e = hipMemcpyAsync(pl->djobs, hj.data(), sizeof(BJob) * count, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && !items.empty())
        e = hipMemcpyAsync(pl->ditems, items.data(), sizeof(BItem) * items.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        int st = hip_check(ctx, e, "batched gemm plan");
        batched_gemm_destroy(ctx, pl);
        return st;
    }
    // persistent grids are a multiple of the CU count (launch_class)
    pl->grid = ctx->num_cus > 0 ? ctx->num_cus : 1;
    // capture [counter reset, one persistent kernel per class] in a graph
    e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        k_queue_reset<<<1, 128, 0, ctx->stream>>>(pl->dcounter);
        for (int c = 0; c < NCLASS; ++c) {
            if (kind == EK_F64) launch_class<double>(c, pl, ctx->stream);
            else launch_class<float>(c, pl, ctx->stream);
        }
        e = hipStreamEndCapture(ctx->stream, &pl->graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        int st = hip_check(ctx, e, "batched gemm graph capture");
        batched_gemm_destroy(ctx, pl);
        return st;
    }
    *out = pl;
    return MODP_OK;
}

}  // namespace

int batched_gemm_run(modp_ctx *ctx, modp_gemm_plan *pl)
{
    if (pl->items == 0) return MODP_OK;
    TimedLaunch tl(ctx, "batched_gemm");
    return hip_check(ctx, hipGraphLaunch(pl->exec, ctx->stream), "hipGraphLaunch");
}

int batched_gemm_destroy(modp_ctx *ctx, modp_gemm_plan *pl)
{
    if (!pl) return MODP_OK;
    (void)hipStreamSynchronize(ctx->stream);
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    if (pl->djobs) (void)hipFree(pl->djobs);
    if (pl->ditems) (void)hipFree(pl->ditems);
    if (pl->dcounter) (void)hipFree(pl->dcounter);
    delete pl;
    return MODP_OK;
}

// ---------------------------------------------------------------------------
// batched elimination: one wavefront per instance, the matrix in LDS
// ---------------------------------------------------------------------------

struct EJob {
    const void *A; int64_t lda;
    void *U; int64_t ldu;
    void *L; int64_t ldl;
    int64_t *row_info;
    int64_t *col_info;
    int64_t *counts;
    int m, n;
    int pitch;  // LDS row pitch in elements
    int cls;    // LDS element: 0 = 8-bit (p < 2^8), 1 = 16-bit (p < 2^16), 2 = 32-bit
    ModP mp;
    const uint16_t *itab;  // inverses mod p (prime p < 2^16), or null: extended Euclid
};

// Instances are grouped by the LDS element their modulus needs (8-bit
// residues below 2^8, 16-bit below 2^16, else 32-bit): each class has its own
// kernel, queue, LDS slot (sized for its largest instance) and persistent
// grid, so the small-element classes keep their occupancy.
constexpr int ECLASS = 3;
struct ElimClass {
    int64_t off = 0, count = 0;  // range of the job table
    size_t slot = 0;             // LDS bytes per workgroup (one wavefront)
    int bk_off = 0;              // offset of the pivot bookkeeping in the slot
    int grid = 0;
};

struct modp_elim_plan {
    int op = 0, kind = EK_F64;
    int64_t count = 0;
    EJob *djobs = nullptr;
    int *dcounter = nullptr;  // one queue counter per class, 64 bytes apart
    ElimClass cls[ECLASS];
    uint16_t *dtab = nullptr;  // inverse tables of the distinct primes below 2^16
    hipStream_t side[ECLASS - 1] = {nullptr, nullptr};  // fork streams of the capture
    hipEvent_t fork = nullptr, join[ECLASS - 1] = {nullptr, nullptr};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

constexpr int EMAX = 128;
// pivot values and inverses, pivot columns, row swaps, permutation, and the
// inverse table of an 8-bit prime
constexpr int EBOOK = EMAX * (4 + 4 + 2 + 2 + 2) + 256;

// Every lane of the wavefront works on one instance, so LDS writes by one
// lane are read by another in program order; the compiler must not move LDS
// accesses across these points (the hardware keeps a wavefront's LDS
// operations in order).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}

__global__ void k_elim_counter_reset(int *counter)
{
    if (threadIdx.x < ECLASS && blockIdx.x == 0) counter[16 * threadIdx.x] = 0;
}

// argmax keys: larger key = larger residue, then lower row (rows < 256)
template <typename S> struct EKey {
    typedef uint32_t T;  // residues below 2^16
    static __device__ __forceinline__ T make(uint32_t v, int r) { return (v << 8) | (uint32_t)(255 - r); }
    static __device__ __forceinline__ uint32_t val(T k) { return k >> 8; }
    static __device__ __forceinline__ int row(T k) { return 255 - (int)(k & 255u); }
};
template <> struct EKey<uint32_t> {
    typedef uint64_t T;
    static __device__ __forceinline__ T make(uint32_t v, int r) { return ((uint64_t)v << 8) | (uint64_t)(255 - r); }
    static __device__ __forceinline__ uint32_t val(T k) { return (uint32_t)(k >> 8); }
    static __device__ __forceinline__ int row(T k) { return 255 - (int)(k & 255u); }
};

// Row update x - l pr (mod p) of one element.  For p < 2^16 (SMALL) the
// caller passes c = p - pr (mod p), so t = x + l c < p + p^2 < 2^32 is
// congruent and reduce_small finishes it with 24-bit integer multiplies.
// Otherwise c = pr and the FP64 reduction runs (|x - l pr| < 2^52 is exact).
// t mod p for p < 2^16, t < 2^32 and t/p < 2^19: q = trunc(fl(fl(t) fl(1/p)
// + 1/2)) is t/p + 1/2 truncated up to an error below 1/8 (the roundings of
// t and 1/p contribute at most (t/p) 2^-23 < 2^-4, the fused rounding at most
// half an ulp of a value below 2^19, 2^-5), so r = t - q p lies in
// (-5p/8, 5p/8) and one fix-up, min(r, r + p) in unsigned arithmetic, brings
// it into [0, p).
__device__ __forceinline__ uint32_t reduce_small(uint32_t t, const ModP &mp)
{
    const uint32_t q = (uint32_t)fmaf((float)t, mp.pinvf, 0.5f);
    const uint32_t r = t - __umul24(q, mp.p32);
    return min(r, r + mp.p32);
}

template <bool SMALL>
__device__ __forceinline__ uint32_t row_update(uint32_t x, uint32_t l, uint32_t c, const ModP &mp)
{
    if (SMALL) return reduce_small(__umul24(l, c) + x, mp);
    return (uint32_t)dmod(fma(-(double)l, (double)c, (double)x), mp);
}

// a b mod p without the run-time modulus-size branch of mulmod
template <bool SMALL> __device__ __forceinline__ uint32_t mulmod_c(uint32_t a, uint32_t b, const ModP &mp)
{
    return SMALL ? row_update<true>(0u, a, b, mp) : (uint32_t)dmod((double)a * (double)b, mp);
}

// ---------------------------------------------------------------------------
// Blocked elimination of one instance.  Pivots are taken one at a time, but
// each pivot's update is applied at once only to the columns of the current
// panel (KB columns from the panel's first); the trailing columns receive the
// panel's pivots together when the panel ends ("flush"):
//   pivot rows, in order: V_j = (x_j - sum_{i<j} L_{r_j,i} V_i) inv_j
//   other rows:           x_r - sum_j L_{r,j} V_j
// with one modular reduction per element per panel instead of one per pivot
// (delayed reduction: below 2^8 the sum stays under 2^20 in 32-bit integers,
// below 2^16 under 2^36 in FP64; 32-bit residues flush after every pivot).
// L_{r,j} is the entry of row r in pivot column j when pivot j was taken,
// which the panel's immediate updates keep current; for rref those entries
// are cleared after the flush (the rows above the pivot are reduced too).
// A zero column under pluq's rule flushes first, so the rest-of-block test
// and the column swap see current values.
// ---------------------------------------------------------------------------

template <int ACC> struct Acc;
template <> struct Acc<0> {  // residues < 2^8, 32-bit integer sums
    typedef uint32_t V;      // stored: p - v (mod p), added
    static __device__ __forceinline__ V from(uint32_t v, const ModP &mp) { return v ? mp.p32 - v : 0u; }
    typedef uint32_t S;
    static __device__ __forceinline__ S init(uint32_t x) { return x; }
    static __device__ __forceinline__ S madd(S a, uint32_t l, V v) { return a + __umul24(l, v); }
    static __device__ __forceinline__ uint32_t reduce(S a, const ModP &mp) { return reduce_small(a, mp); }
};
template <> struct Acc<1> {  // FP64 sums, exact below 2^53
    typedef double V;        // stored: v, subtracted
    static __device__ __forceinline__ V from(uint32_t v, const ModP &) { return (double)v; }
    typedef double S;
    static __device__ __forceinline__ S init(uint32_t x) { return (double)x; }
    static __device__ __forceinline__ S madd(S a, uint32_t l, V v) { return fma(-(double)l, v, a); }
    static __device__ __forceinline__ uint32_t reduce(S a, const ModP &mp) { return (uint32_t)dmod(a, mp); }
};

// Rows r0..r1-1 (all below 64, or all from 64 on: HALF picks the multiplier
// register) of the trailing update x_r - sum_j L_{r,j} V_j.  Pivots past
// the panel's count carry zero multipliers and values, so every row runs
// the same KB multiply-adds; RB rows' loads are issued before their stores.
template <typename S, int KB, int ACC, bool WIDE, int RB>
__device__ __forceinline__ void flush_rows(S *M, int pitch, int r0, int r1, int c0, int c1, const uint32_t *lq,
                                           const typename Acc<ACC>::V *v0, const typename Acc<ACC>::V *v1,
                                           const ModP &mp)
{
    typedef Acc<ACC> A;
    int r = r0;
    for (; r + RB <= r1; r += RB) {
        uint32_t x0[RB], x1[RB];
#pragma unroll
        for (int t = 0; t < RB; ++t) {
            x0[t] = M[(r + t) * pitch + c0];
            x1[t] = WIDE ? (uint32_t)M[(r + t) * pitch + c1] : 0u;
        }
#pragma unroll
        for (int t = 0; t < RB; ++t) {
            typename A::S a0 = A::init(x0[t]), a1 = A::init(x1[t]);
#pragma unroll
            for (int jj = 0; jj < KB; ++jj) {
                const uint32_t l = (uint32_t)__builtin_amdgcn_readlane((int)lq[jj], (r + t) & 63);
                a0 = A::madd(a0, l, v0[jj]);
                if (WIDE) a1 = A::madd(a1, l, v1[jj]);
            }
            M[(r + t) * pitch + c0] = (S)A::reduce(a0, mp);
            if (WIDE) M[(r + t) * pitch + c1] = (S)A::reduce(a1, mp);
        }
    }
    for (; r < r1; ++r) {
        typename A::S a0 = A::init(M[r * pitch + c0]), a1 = A::init(WIDE ? (uint32_t)M[r * pitch + c1] : 0u);
#pragma unroll
        for (int jj = 0; jj < KB; ++jj) {
            const uint32_t l = (uint32_t)__builtin_amdgcn_readlane((int)lq[jj], r & 63);
            a0 = A::madd(a0, l, v0[jj]);
            if (WIDE) a1 = A::madd(a1, l, v1[jj]);
        }
        M[r * pitch + c0] = (S)A::reduce(a0, mp);
        if (WIDE) M[r * pitch + c1] = (S)A::reduce(a1, mp);
    }
}

typedef int ev4i __attribute__((ext_vector_type(4)));
typedef int ev16i __attribute__((ext_vector_type(16)));

// Rows a..b-1 of the trailing update for 8-bit residues on int8 MFMA
// (v_mfma_i32_32x32x32_i8), one 32 x 32 tile of rows x trailing columns per
// instruction: D = x + sum_j L[r][j] W[j][c] with L the multipliers (pivot
// columns pc[j]) and W = -V (the pivot rows' trailing values, written by the
// flush's first step), both as balanced residues in (-p/2, p/2] (|.| <= 125),
// so |sum| <= 8 * 125^2 and D fits int32.  Pivots past np contribute zeros;
// K bytes 8..31 of the instruction are zero.  Operand layouts (lane l): A row
// l % 32 and B column l % 32 with K bytes 16 (l / 32) .. +15; accumulator
// column l % 32, rows 8 (e / 4) + 4 (l / 32) + e % 4.  Rows and columns past
// the range repeat the last one (same inputs, same value stored).
// D mod p: q = rint(fl(D) fl(1/p)) is within 0.01 of D/p (|D|/p < 2^16), so
// r = D - q p lies in (-0.51 p, 0.51 p) and one fix-up, min(r, r + p) in
// unsigned arithmetic, finishes it.
template <int KB>
__device__ __forceinline__ void flush_rows_mfma8(uint8_t *M, int pitch, int a, int b, int pe, int n, int prow,
                                                 int np, const int *pc, const ModP &mp)
{
    if (a >= b) return;
    const int lane = threadIdx.x, l32 = lane & 31, hi = lane >> 5;
    const int p = (int)mp.p32, half = p >> 1;
    const int ntc = (n - pe + 31) >> 5;  // column tiles, at most 4
    ev4i bop[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        uint32_t w[2] = {0u, 0u};
        if (ct < ntc) {
            const int c = min(pe + 32 * ct + l32, n - 1);
#pragma unroll
            for (int jj = 0; jj < KB; ++jj) {
                if (jj < np) {
                    const int v = M[(prow + jj) * pitch + c];
                    const int nv = v > half ? p - v : -v;  // -v, balanced
                    w[jj >> 2] |= ((uint32_t)nv & 255u) << (8 * (jj & 3));
                }
            }
        }
        bop[ct] = hi ? ev4i{0, 0, 0, 0} : ev4i{(int)w[0], (int)w[1], 0, 0};
    }
    for (int r0 = a; r0 < b; r0 += 32) {
        uint32_t w[2] = {0u, 0u};
        const int rr = min(r0 + l32, b - 1);
#pragma unroll
        for (int jj = 0; jj < KB; ++jj) {
            if (jj < np) {
                const int v = M[rr * pitch + pc[jj]];
                const int bv = v > half ? v - p : v;  // balanced
                w[jj >> 2] |= ((uint32_t)bv & 255u) << (8 * (jj & 3));
            }
        }
        const ev4i aop = hi ? ev4i{0, 0, 0, 0} : ev4i{(int)w[0], (int)w[1], 0, 0};
        int ro[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) ro[e] = min(r0 + 8 * (e >> 2) + 4 * hi + (e & 3), b - 1) * pitch;
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            if (ct < ntc) {
                const int c = min(pe + 32 * ct + l32, n - 1);
                ev16i acc;
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[e] = M[ro[e] + c];
                acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(aop, bop[ct], acc, 0, 0, 0);
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    // q = round(d / p): adding 1.5 2^23 leaves the rounded
                    // quotient in the low mantissa bits (|q| < 2^17)
                    const int d = acc[e];
                    const int q = __float_as_int(fmaf((float)d, mp.pinvf, 12582912.0f)) - 0x4B400000;
                    const uint32_t r = (uint32_t)(d - __mul24(q, p));
                    M[ro[e] + c] = (uint8_t)min(r, r + (uint32_t)p);
                }
            }
        }
    }
}

// Rows a..b-1 of the trailing update for 16-bit residues on FP64 MFMA
// (v_mfma_f64_16x16x4f64, two per 16 x 16 tile for the eight pivots):
// D = x - sum_j L[r][j] V[j][c] with |sum| < 8 2^32, exact in FP64, then
// reduced mod p.  Operand layouts (lane l): A row l & 15 and B column l & 15
// with K index l >> 4; accumulator column l & 15, rows (l >> 4) + 4 e.  Rows
// and columns past the range repeat the last one.  V comes from the pivot
// rows' trailing columns, written by the flush's first step.
template <int KB>
__device__ __forceinline__ void flush_rows_mfma16(uint16_t *M, int pitch, int a, int b, int pe, int n, int prow,
                                                  int np, const int16_t *pivcol, const ModP &mp)
{
    static_assert(KB == 8, "two K=4 steps per tile");
    if (a >= b) return;
    const int lane = threadIdx.x, l16 = lane & 15, kq = lane >> 4;
    const int ntc = (n - pe + 15) >> 4;  // column tiles, at most 8
    // this lane's pivots kq and kq + 4 (their columns from the bookkeeping)
    const bool ka = kq < np, kb = kq + 4 < np;
    const int pca = ka ? (int)pivcol[prow + kq] : 0;
    const int pcb = kb ? (int)pivcol[prow + kq + 4] : 0;
    const int va = (prow + (ka ? kq : 0)) * pitch, vb = (prow + (kb ? kq + 4 : 0)) * pitch;
    for (int r0 = a; r0 < b; r0 += 16) {
        const int rr = min(r0 + l16, b - 1);
        const double a0 = ka ? (double)M[rr * pitch + pca] : 0.0;
        const double a1 = kb ? (double)M[rr * pitch + pcb] : 0.0;
        int ro[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) ro[e] = min(r0 + kq + 4 * e, b - 1) * pitch;
        // G column tiles at a time: all their loads, then the products, then
        // the reductions and stores (tiles past the last repeat it)
        auto tiles = [&](int ct0, auto G) {
            constexpr int NG = decltype(G)::value;
            v4d acc[NG];
            double b0[NG], b1[NG];
            int cc[NG];
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                cc[g] = min(pe + 16 * (ct0 + g) + l16, n - 1);
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[g][e] = (double)M[ro[e] + cc[g]];
                b0[g] = ka ? -(double)M[va + cc[g]] : 0.0;
                b1[g] = kb ? -(double)M[vb + cc[g]] : 0.0;
            }
#pragma unroll
            for (int g = 0; g < NG; ++g) acc[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0[g], acc[g], 0, 0, 0);
#pragma unroll
            for (int g = 0; g < NG; ++g) acc[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1[g], acc[g], 0, 0, 0);
#pragma unroll
            for (int g = 0; g < NG; ++g)
#pragma unroll
                for (int e = 0; e < 4; ++e) M[ro[e] + cc[g]] = (uint16_t)dmod(acc[g][e], mp);
        };
        int ct = 0;
        for (; ct + 4 <= ntc; ct += 4) tiles(ct, std::integral_constant<int, 4>());
        for (; ct < ntc; ++ct) tiles(ct, std::integral_constant<int, 1>());
    }
}

// Trailing columns pe..n-1 of the pending pivots (rows prow..prow+np-1) on
// lanes over columns: lane q owns columns pe+q and (WIDE) pe+64+q; lanes
// past the last column repeat it, computing and storing the same value.
template <typename S, int KB, int ACC, bool WIDE>
__device__ __forceinline__ void flush_cols(S *M, int pitch, int m, int n, int op, int prow, int np, int pe,
                                           const int *pc, const int16_t *pivcol, const uint32_t *inv,
                                           const uint32_t *lq0, const uint32_t *lq1, const ModP &mp)
{
    typedef Acc<ACC> A;
    constexpr bool SMALL = sizeof(S) < 4;
    const int lane = threadIdx.x;
    const int c0 = min(pe + lane, n - 1), c1 = min(pe + 64 + lane, n - 1);
    typename A::V v0[KB], v1[KB];
    // pivot rows, in order: V_j = (x_j - sum_{i<j} L_{r_j,i} V_i) inv_j
#pragma unroll
    for (int jj = 0; jj < KB; ++jj) {
        v0[jj] = v1[jj] = typename A::V(0);
        if (jj < np) {
            const int r = prow + jj;
            typename A::S a0 = A::init(M[r * pitch + c0]), a1 = A::init(WIDE ? (uint32_t)M[r * pitch + c1] : 0u);
#pragma unroll
            for (int ii = 0; ii < jj; ++ii) {
                const uint32_t l = (uint32_t)__builtin_amdgcn_readlane((int)(r < 64 ? lq0[ii] : lq1[ii]), r & 63);
                a0 = A::madd(a0, l, v0[ii]);
                if (WIDE) a1 = A::madd(a1, l, v1[ii]);
            }
            const uint32_t y0 = mulmod_c<SMALL>(A::reduce(a0, mp), inv[jj], mp);
            M[r * pitch + c0] = (S)y0;
            v0[jj] = A::from(y0, mp);
            if (WIDE) {
                const uint32_t y1 = mulmod_c<SMALL>(A::reduce(a1, mp), inv[jj], mp);
                M[r * pitch + c1] = (S)y1;
                v1[jj] = A::from(y1, mp);
            }
        }
    }
    wave_sync();
    // every other row
    constexpr int RB = 4;
    auto rows = [&](int a, int b) {
        if (a < min(b, 64)) flush_rows<S, KB, ACC, WIDE, RB>(M, pitch, a, min(b, 64), c0, c1, lq0, v0, v1, mp);
        if (max(a, 64) < b) flush_rows<S, KB, ACC, WIDE, RB>(M, pitch, max(a, 64), b, c0, c1, lq1, v0, v1, mp);
    };
    if (sizeof(S) == 1) {
        uint8_t *M8 = reinterpret_cast<uint8_t *>(M);
        flush_rows_mfma8<KB>(M8, pitch, prow + np, m, pe, n, prow, np, pc, mp);
        if (op == MODP_ELIM_RREF) flush_rows_mfma8<KB>(M8, pitch, 0, prow, pe, n, prow, np, pc, mp);
    } else if (sizeof(S) == 2 && KB == 8) {
        uint16_t *M16 = reinterpret_cast<uint16_t *>(M);
        flush_rows_mfma16<8>(M16, pitch, prow + np, m, pe, n, prow, np, pivcol, mp);
        if (op == MODP_ELIM_RREF) flush_rows_mfma16<8>(M16, pitch, 0, prow, pe, n, prow, np, pivcol, mp);
    } else {
        rows(prow + np, m);
        if (op == MODP_ELIM_RREF) rows(0, prow);
    }
    if (op == MODP_ELIM_RREF) {
        // the panel's pivot rows take the later pivots of the panel
#pragma unroll
        for (int own = 0; own + 1 < KB; ++own) {
            if (own + 1 < np) {
                const int r = prow + own;
                typename A::S a0 = A::init(M[r * pitch + c0]), a1 = A::init(WIDE ? (uint32_t)M[r * pitch + c1] : 0u);
#pragma unroll
                for (int jj = own + 1; jj < KB; ++jj) {
                    const uint32_t l = (uint32_t)__builtin_amdgcn_readlane((int)(r < 64 ? lq0[jj] : lq1[jj]), r & 63);
                    a0 = A::madd(a0, l, v0[jj]);
                    if (WIDE) a1 = A::madd(a1, l, v1[jj]);
                }
                M[r * pitch + c0] = (S)A::reduce(a0, mp);
                if (WIDE) M[r * pitch + c1] = (S)A::reduce(a1, mp);
            }
        }
    }
    wave_sync();
}

template <typename S, int KB, int ACC>
__device__ __forceinline__ void flush(S *M, int pitch, int m, int n, int op, int prow, int np, int pe,
                                      const int16_t *pivcol, const uint32_t *pivinv, const ModP &mp)
{
    if (np <= 0) return;
    const int lane = threadIdx.x;
    int pc[KB];
    uint32_t inv[KB], lq0[KB], lq1[KB];
#pragma unroll
    for (int jj = 0; jj < KB; ++jj) {
        pc[jj] = 0;
        inv[jj] = lq0[jj] = lq1[jj] = 0u;
        if (jj < np) {
            pc[jj] = __builtin_amdgcn_readfirstlane((int)pivcol[prow + jj]);
            inv[jj] = (uint32_t)__builtin_amdgcn_readfirstlane((int)pivinv[prow + jj]);
            lq0[jj] = lane < m ? (uint32_t)M[lane * pitch + pc[jj]] : 0u;
            lq1[jj] = lane + 64 < m ? (uint32_t)M[(lane + 64) * pitch + pc[jj]] : 0u;
        }
    }
    if (pe < n) {
        if (n - pe > 64)
            flush_cols<S, KB, ACC, true>(M, pitch, m, n, op, prow, np, pe, pc, pivcol, inv, lq0, lq1, mp);
        else
            flush_cols<S, KB, ACC, false>(M, pitch, m, n, op, prow, np, pe, pc, pivcol, inv, lq0, lq1, mp);
    }
    if (op == MODP_ELIM_RREF) {
        // the multipliers have been applied: clear the pivot columns but the pivots
#pragma unroll
        for (int jj = 0; jj < KB; ++jj)
            if (jj < np)
                for (int r = lane; r < m; r += 64)
                    if (r != prow + jj) M[r * pitch + pc[jj]] = S(0);
        wave_sync();
    }
}

// One instance on one wavefront (64 threads): argmax over rows with the lanes
// over rows (DPP reduction of (residue, -row) keys), row swaps and pivot-row
// normalisation with the lanes over columns, the panel's immediate updates
// with the lanes over rows and the flushes with the lanes over columns.
// S is the LDS element; every control value (m, n, row, col, the pivot) is
// wave-uniform, so the branches are scalar.
template <typename T, typename S, int KB, int ACC>
__device__ __forceinline__ void elim_instance(const EJob &j, int op, unsigned char *esm, int bk_off)
{
    typedef EKey<S> K;
    typedef typename K::T key_t;
    constexpr bool SMALL = sizeof(S) < 4;
    S *M = reinterpret_cast<S *>(esm);
    uint32_t *pivval = reinterpret_cast<uint32_t *>(esm + bk_off);
    uint32_t *pivinv = pivval + EMAX;
    int16_t *pivcol = reinterpret_cast<int16_t *>(pivinv + EMAX);
    int16_t *ipiv = pivcol + EMAX;
    int16_t *perm = ipiv + EMAX;
    uint8_t *itab8 = reinterpret_cast<uint8_t *>(perm + EMAX);
    const int lane = threadIdx.x;
    const int m = j.m, n = j.n, pitch = j.pitch;
    const ModP &mp = j.mp;
    const T *A = (const T *)j.A;
    // 8-bit primes: the inverse table goes to LDS (one lookup per pivot);
    // 16-bit primes look theirs up in global memory (L2-resident: one table
    // per prime serves the whole batch)
    const bool tab8 = sizeof(S) == 1 && j.itab != nullptr;
    if (tab8) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int e = 4 * lane + t;
            itab8[e] = (uint8_t)(e < (int)mp.p32 ? j.itab[e] : 0u);
        }
    }
    // column-major input, lanes over rows (two rows per lane at most):
    // 16 columns of loads in flight before their LDS stores.  Lanes past the
    // last row or column repeat it, loading and storing the same value as
    // its owner, so no lane is masked off.
    {
        const int r0 = min(lane, m - 1), r1 = min(lane + 64, m - 1);
        const bool two = m > 64;
        for (int c0 = 0; c0 < n; c0 += 16) {
            T x0[16], x1[16];
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int64_t c = min(c0 + t, n - 1);
                x0[t] = A[r0 + c * j.lda];
                x1[t] = two ? A[r1 + c * j.lda] : T(0);
            }
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int c = min(c0 + t, n - 1);
                M[r0 * pitch + c] = (S)to_res<T>(x0[t], mp);
                if (two) M[r1 * pitch + c] = (S)to_res<T>(x1[t], mp);
            }
        }
    }
    wave_sync();
    int row = 0, col = 0, last = n - 1, nr = 0, nc = 0;
    int prow = 0;               // first pivot row whose trailing update is pending
    int pe = KB < n ? KB : n;   // end of the current panel's columns
    bool bad = false, ended = false;
    // the panel update computes the next pivot column's keys on the way
    key_t pre = 0;
    bool have_pre = false;
    while (row < m && col < n) {
        // largest residue of column col in rows row..m-1, lowest row on ties
        key_t mine = 0;
        if (have_pre) {
            mine = pre;
            have_pre = false;
        } else {
            if (lane >= row && lane < m) mine = K::make(M[lane * pitch + col], lane);
            if (lane + 64 >= row && lane + 64 < m) {
                const key_t k = K::make(M[(lane + 64) * pitch + col], lane + 64);
                mine = k > mine ? k : mine;
            }
        }
        const key_t best = wave_max(mine);
        const uint32_t bv = K::val(best);
        if (bv == 0 && op == MODP_ELIM_PLUQ) {
            flush<S, KB, ACC>(M, pitch, m, n, op, prow, row - prow, pe, pivcol, pivinv, mp);
            prow = row;
            // is the whole block rows row.., columns col.. zero?
            int nz = 0;
            for (int c = col + lane; c < n; c += 64)
                for (int r = row; r < m; ++r) nz |= M[r * pitch + c] != 0;
            if (!__any(nz)) {
                // the reference's tail: (c, n) for c = col..n-1 in order,
                // which rotates columns col..n-1 right by one
                const int ca = col + 1 + lane, cb = col + 65 + lane;
                for (int r = 0; r < m; ++r) {
                    const S first = M[r * pitch + n - 1];
                    const S va = ca < n ? M[r * pitch + ca - 1] : S(0);
                    const S vb = cb < n ? M[r * pitch + cb - 1] : S(0);
                    wave_sync();
                    if (ca < n) M[r * pitch + ca] = va;
                    if (cb < n) M[r * pitch + cb] = vb;
                    if (lane == 0) M[r * pitch + col] = first;
                    wave_sync();
                }
                if (lane == 0)
                    for (int c = col, k = nc; c < n; ++c, ++k) {
                        j.col_info[2 * k] = c + 1;
                        j.col_info[2 * k + 1] = n;
                    }
                nc += n - col;
                ended = true;
                break;
            }
            // swap with the last column not yet swapped; retry this column
            for (int r = lane; r < m; r += 64) {
                const S t = M[r * pitch + col];
                M[r * pitch + col] = M[r * pitch + last];
                M[r * pitch + last] = t;
            }
            if (lane == 0) {
                j.col_info[2 * nc] = col + 1;
                j.col_info[2 * nc + 1] = last + 1;
            }
            ++nc;
            --last;
            wave_sync();
            continue;
        }
        if (bv == 0) {
            ++col;  // lu, rref: no pivot in this column
        } else {
            const uint32_t inv = tab8 ? (uint32_t)itab8[bv]
                                      : (sizeof(S) == 2 && j.itab ? (uint32_t)j.itab[bv] : modinv_fast(bv, mp));
            if (inv == 0) { bad = true; break; }
            const int I = K::row(best);
            // swap rows row <-> I (every column) and normalise the pivot row
            // over the panel (the flush normalises the trailing columns)
            {
                const int ca = min(lane, n - 1), cb = min(lane + 64, n - 1);
                const uint32_t a0 = M[row * pitch + ca], b0 = M[I * pitch + ca];
                const uint32_t a1 = M[row * pitch + cb], b1 = M[I * pitch + cb];
                wave_sync();
                M[I * pitch + ca] = (S)a0;
                M[row * pitch + ca] = (S)((ca >= col && ca < pe) ? mulmod_c<SMALL>(b0, inv, mp) : b0);
                if (n > 64) {
                    M[I * pitch + cb] = (S)a1;
                    M[row * pitch + cb] = (S)((cb >= col && cb < pe) ? mulmod_c<SMALL>(b1, inv, mp) : b1);
                }
            }
            if (lane == 0) {
                pivval[row] = bv;
                pivinv[row] = inv;
                pivcol[row] = (int16_t)col;
                ipiv[row] = (int16_t)I;
                if (op == MODP_ELIM_PLUQ && I != row) {
                    j.row_info[2 * nr] = row + 1;
                    j.row_info[2 * nr + 1] = I + 1;
                }
            }
            if (op == MODP_ELIM_PLUQ && I != row) ++nr;
            wave_sync();
            // the panel's columns col+1..pe-1, lanes over rows: rows below the
            // pivot (pluq, lu) or every other row (rref)
            if (col + 1 < pe) {
                // lane q owns rows q and q+64; lanes past the last row repeat
                // it (same multiplier, same values, same stores).  All loads
                // are issued before the updates; reads past column pe-1 stay
                // inside the slot (the bookkeeping follows the matrix), only
                // the panel is written.
                const int i0 = op == MODP_ELIM_RREF ? 0 : row + 1;
                const int ra = min(lane, m - 1), rb = min(lane + 64, m - 1);
                const bool two = m > 64;
                uint32_t kp[KB], xa[KB], xb[KB];
#pragma unroll
                for (int t = 0; t + 1 < KB; ++t) {
                    const uint32_t v = M[row * pitch + col + 1 + t];
                    kp[t] = SMALL ? (v ? mp.p32 - v : 0u) : v;
                    xa[t] = M[ra * pitch + col + 1 + t];
                    xb[t] = two ? (uint32_t)M[rb * pitch + col + 1 + t] : 0u;
                }
                const uint32_t la = (ra >= i0 && ra != row) ? (uint32_t)M[ra * pitch + col] : 0u;
                const uint32_t lb = (two && rb >= i0 && rb != row) ? (uint32_t)M[rb * pitch + col] : 0u;
                wave_sync();
#pragma unroll
                for (int t = 0; t + 1 < KB; ++t) {
                    if (col + 1 + t < pe) {
                        const uint32_t ya = row_update<SMALL>(xa[t], la, kp[t], mp);
                        const uint32_t yb = two ? row_update<SMALL>(xb[t], lb, kp[t], mp) : 0u;
                        M[ra * pitch + col + 1 + t] = (S)ya;
                        if (two) M[rb * pitch + col + 1 + t] = (S)yb;
                        if (t == 0) {
                            // keys of column col+1 over rows row+1.. for the next argmax
                            pre = ra > row ? K::make(ya, ra) : key_t(0);
                            if (two && rb > row) {
                                const key_t kb = K::make(yb, rb);
                                pre = kb > pre ? kb : pre;
                            }
                            have_pre = true;
                        }
                    }
                }
                wave_sync();
            }
            ++row;
            ++col;
        }
        if (col >= pe || row >= m) {
            flush<S, KB, ACC>(M, pitch, m, n, op, prow, row - prow, pe, pivcol, pivinv, mp);
            prow = row;
            pe = col + KB < n ? col + KB : n;
        }
    }
    if (!bad && !ended && prow < row) flush<S, KB, ACC>(M, pitch, m, n, op, prow, row - prow, pe, pivcol, pivinv, mp);
    const int rank = row;
    // outputs, column-major: lanes over rows write consecutive elements,
    // 16 columns of LDS reads before their stores; lanes past the last row
    // or column repeat it (same address, same value)
    const int r0 = min(lane, m - 1), r1 = min(lane + 64, m - 1);
    const bool two = m > 64;
    {
        T *U = (T *)j.U;
        // first column of row r's U part (n: none)
        const int f0 = r0 < rank ? (op == MODP_ELIM_RREF ? 0 : (int)pivcol[r0]) : n;
        const int f1 = r1 < rank ? (op == MODP_ELIM_RREF ? 0 : (int)pivcol[r1]) : n;
        for (int c0 = 0; c0 < n; c0 += 16) {
            uint32_t x0[16], x1[16];
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int c = min(c0 + t, n - 1);
                x0[t] = M[r0 * pitch + c];
                x1[t] = two ? (uint32_t)M[r1 * pitch + c] : 0u;
            }
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int c = min(c0 + t, n - 1);
                U[r0 + (int64_t)c * j.ldu] = from_res<T>(c >= f0 ? x0[t] : 0u);
                if (two) U[r1 + (int64_t)c * j.ldu] = from_res<T>(c >= f1 ? x1[t] : 0u);
            }
        }
    }
    if (op != MODP_ELIM_RREF) {
        // L: column k < rank holds the pivot value on the diagonal and the
        // multipliers (pivot column k) below it
        T *L = (T *)j.L;
        for (int k0 = 0; k0 < m; k0 += 16) {
            int pc[16];
            uint32_t pv[16], x0[16], x1[16];
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int kk = min(min(k0 + t, m - 1), rank > 0 ? rank - 1 : 0);
                pc[t] = rank > 0 ? (int)pivcol[kk] : 0;
                pv[t] = rank > 0 ? pivval[kk] : 0u;
            }
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                x0[t] = M[r0 * pitch + pc[t]];
                x1[t] = two ? (uint32_t)M[r1 * pitch + pc[t]] : 0u;
            }
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int k = min(k0 + t, m - 1);
                const uint32_t v0 = k < rank ? (r0 == k ? pv[t] : (r0 > k ? x0[t] : 0u)) : 0u;
                L[r0 + (int64_t)k * j.ldl] = from_res<T>(v0);
                if (two) {
                    const uint32_t v1 = k < rank ? (r1 == k ? pv[t] : (r1 > k ? x1[t] : 0u)) : 0u;
                    L[r1 + (int64_t)k * j.ldl] = from_res<T>(v1);
                }
            }
        }
    }
    if (op == MODP_ELIM_LU) {
        // row permutation: the identity with the pivot swaps applied in order
        // (built in LDS, written by all lanes)
        if (lane == 0) {
            for (int q = 0; q < m; ++q) perm[q] = (int16_t)q;
            for (int q = 0; q < rank; ++q) {
                const int16_t t = perm[q];
                perm[q] = perm[ipiv[q]];
                perm[ipiv[q]] = t;
            }
        }
        wave_sync();
        for (int q = lane; q < m; q += 64) j.row_info[q] = perm[q] + 1;
        for (int q = lane; q < rank; q += 64) j.col_info[q] = pivcol[q] + 1;
    } else if (op == MODP_ELIM_RREF) {
        for (int q = lane; q < rank; q += 64) j.row_info[q] = pivcol[q] + 1;
    }
    if (lane == 0) {
        j.counts[0] = op == MODP_ELIM_PLUQ ? nr : 0;
        j.counts[1] = op == MODP_ELIM_PLUQ ? nc : 0;
        j.counts[2] = bad ? -1 : rank;
    }
    wave_sync();
}

// Persistent kernel of one class: workgroup w (one wavefront) starts with
// instance w of the class (sorted by cost, largest first; the grid never
// exceeds the count), then takes instances grid, grid+1, ... from the
// class's queue as it finishes.  Taking the next index only after an
// instance is done keeps a workgroup that starts early from claiming two
// large instances while later workgroups are still being dispatched.
template <typename T, int CLS>
__global__ __launch_bounds__(64) void k_elim_wave(const EJob *__restrict__ jobs, int64_t njobs, int op,
                                                  int *__restrict__ counter, int bk_off)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char esm[];
    int64_t it = blockIdx.x;
    while (it < njobs) {
        const EJob j = jobs[it];
        if (CLS == 0) elim_instance<T, uint8_t, 8, 0>(j, op, esm, bk_off);
        else if (CLS == 1) elim_instance<T, uint16_t, 8, 1>(j, op, esm, bk_off);
        else elim_instance<T, uint32_t, 1, 1>(j, op, esm, bk_off);
        int next = 0;
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        it = (int64_t)gridDim.x + __builtin_amdgcn_readfirstlane(next);
    }
}

template <typename T, int CLS> hipError_t class_kernel(const modp_elim_plan *pl, hipStream_t st, bool allow)
{
    const ElimClass &ec = pl->cls[CLS];
    // the dynamic LDS can exceed the default 64 KiB (32-bit residues of
    // 128 x 128 instances)
    if (allow)
        return hipFuncSetAttribute((const void *)k_elim_wave<T, CLS>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)ec.slot);
    k_elim_wave<T, CLS><<<(unsigned)ec.grid, 64, ec.slot, st>>>(pl->djobs + ec.off, ec.count, pl->op,
                                                                 pl->dcounter + 16 * CLS, ec.bk_off);
    return hipGetLastError();
}

template <typename T> hipError_t class_launch(const modp_elim_plan *pl, int c, hipStream_t st, bool allow)
{
    if (c == 0) return class_kernel<T, 0>(pl, st, allow);
    if (c == 1) return class_kernel<T, 1>(pl, st, allow);
    return class_kernel<T, 2>(pl, st, allow);
}

int elim_class(int64_t p) { return p < 256 ? 0 : (p < 65536 ? 1 : 2); }

// a row pitch of an odd number of 32-bit words, so the lanes of a column
// walk hit distinct banks
int elim_pitch(int n, int eb)
{
    int pitch = n;
    while ((pitch * eb) % 4 != 0 || ((pitch * eb) / 4) % 2 == 0) ++pitch;
    return pitch;
}

}  // namespace

int batched_elim_destroy(modp_ctx *ctx, modp_elim_plan *pl)
{
    if (!pl) return MODP_OK;
    (void)hipStreamSynchronize(ctx->stream);
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    for (int i = 0; i < ECLASS - 1; ++i) {
        if (pl->join[i]) (void)hipEventDestroy(pl->join[i]);
        if (pl->side[i]) (void)hipStreamDestroy(pl->side[i]);
    }
    if (pl->fork) (void)hipEventDestroy(pl->fork);
    if (pl->dtab) (void)hipFree(pl->dtab);
    if (pl->djobs) (void)hipFree(pl->djobs);
    if (pl->dcounter) (void)hipFree(pl->dcounter);
    delete pl;
    return MODP_OK;
}

// The plan groups the instances by class, sorts each class by cost (largest
// first, so its queue ends with short ones), gives each instance its row
// pitch, sizes each class's LDS slot for its largest instance, uploads the
// job table once and captures [queue reset, one persistent kernel per class]
// in a hipGraph, the classes on parallel branches (forked streams) so they
// run side by side.
int batched_elim_plan(modp_ctx *ctx, int op, int kind, const modp_elim_job *jobs, int64_t count,
                      modp_elim_plan **out)
{
    if (kind == EK_I32) return set_error(ctx, MODP_ERR_DTYPE, "batched elimination: unsupported element type");
    modp_elim_plan *pl = new modp_elim_plan();
    pl->op = op;
    pl->kind = kind;
    pl->count = count;
    std::vector<EJob> hj;
    hj.reserve((size_t)(count > 0 ? count : 0));
    auto cost = [&](int64_t i) {
        const int64_t a = jobs[i].m, b = jobs[i].n;
        return a * b * (a < b ? a : b);
    };
    // inverse tables for the distinct primes below 2^16 (up to 16 MiB of
    // tables; other moduli keep the extended Euclidean algorithm), by
    // inv(i) = -(p div i) inv(p mod i) mod p
    std::map<int64_t, size_t> tab_off;
    std::vector<uint16_t> htab;
    for (int64_t i = 0; i < count; ++i) {
        const int64_t q = jobs[i].p;
        if (q >= 65536 || q < 2 || tab_off.count(q)) continue;
        bool prime = true;
        for (int64_t d = 2; d * d <= q && prime; ++d) prime = q % d != 0;
        if (!prime || htab.size() + (size_t)q > ((size_t)8 << 20)) continue;
        tab_off[q] = htab.size();
        htab.resize(htab.size() + (size_t)q + (q & 1));  // keep 4-byte alignment
        uint16_t *t = htab.data() + tab_off[q];
        t[0] = 0;
        if (q > 1) t[1] = 1;
        for (int64_t a = 2; a < q; ++a) t[a] = (uint16_t)((q - (q / a) * (int64_t)t[q % a] % q) % q);
    }
    hipError_t e = hipSuccess;
    if (!htab.empty()) {
        e = hipMalloc(&pl->dtab, htab.size() * sizeof(uint16_t));
        if (e == hipSuccess)
            e = hipMemcpyAsync(pl->dtab, htab.data(), htab.size() * sizeof(uint16_t), hipMemcpyHostToDevice,
                               ctx->stream);
    }
    for (int c = 0; c < ECLASS; ++c) {
        std::vector<int64_t> order;
        for (int64_t i = 0; i < count; ++i)
            if (elim_class(jobs[i].p) == c) order.push_back(i);
        std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return cost(a) > cost(b); });
        ElimClass &ec = pl->cls[c];
        ec.off = (int64_t)hj.size();
        ec.count = (int64_t)order.size();
        const int eb = c == 0 ? 1 : (c == 1 ? 2 : 4);
        size_t mat_bytes = 16;
        for (int64_t oi : order) {
            const modp_elim_job &s = jobs[oi];
            EJob d;
            d.A = s.A; d.lda = s.lda; d.U = s.U; d.ldu = s.ldu; d.L = s.L; d.ldl = s.ldl;
            d.row_info = s.row_info; d.col_info = s.col_info; d.counts = s.counts;
            d.m = (int)s.m; d.n = (int)s.n;
            d.cls = c;
            d.pitch = elim_pitch(d.n, eb);
            d.mp = make_modp(s.p);
            const auto ti = tab_off.find(s.p);
            d.itab = ti == tab_off.end() ? nullptr : pl->dtab + ti->second;
            mat_bytes = std::max(mat_bytes, (size_t)d.m * (size_t)d.pitch * (size_t)eb);
            hj.push_back(d);
        }
        ec.bk_off = (int)((mat_bytes + 15) & ~(size_t)15);
        ec.slot = (size_t)ec.bk_off + EBOOK;
        // persistent grid: as many single-wavefront workgroups as the LDS holds
        const size_t lds_cu = 160 * 1024;
        int64_t per_cu = (int64_t)(lds_cu / ec.slot);
        if (per_cu > 16) per_cu = 16;
        if (per_cu < 1) per_cu = 1;
        const int64_t g = per_cu * ctx->num_cus;
        ec.grid = (int)(g < ec.count ? g : ec.count);
    }
    if (e == hipSuccess) e = hipMalloc(&pl->djobs, sizeof(EJob) * (count > 0 ? count : 1));
    if (e == hipSuccess) e = hipMalloc(&pl->dcounter, 256);
    if (e == hipSuccess && count > 0)
        e = hipMemcpyAsync(pl->djobs, hj.data(), sizeof(EJob) * count, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // hj is released on return
    // classes with work, the one with the costliest instance (the longest
    // single-wavefront latency) first: its workgroups are dispatched before
    // the other classes fill the CUs' LDS
    int act[ECLASS], na = 0;
    int64_t top[ECLASS] = {0, 0, 0};
    for (int c = 0; c < ECLASS; ++c) {
        if (pl->cls[c].count == 0) continue;
        const EJob &f = hj[(size_t)pl->cls[c].off];
        top[c] = (int64_t)f.m * f.n * (f.m < f.n ? f.m : f.n) * (c == 2 ? 3 : (c == 1 ? 5 : 4));
        act[na++] = c;
    }
    std::sort(act, act + na, [&](int x, int y) { return top[x] > top[y]; });
    for (int i = 0; i < na && e == hipSuccess; ++i)
        e = kind == EK_F64 ? class_launch<double>(pl, act[i], ctx->stream, true)
                           : class_launch<float>(pl, act[i], ctx->stream, true);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&pl->fork, hipEventDisableTiming);
    for (int i = 0; i + 1 < na && e == hipSuccess; ++i) {
        e = hipStreamCreateWithFlags(&pl->side[i], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&pl->join[i], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t le = hipSuccess;
        k_elim_counter_reset<<<1, 64, 0, ctx->stream>>>(pl->dcounter);
        le = hipGetLastError();
        if (le == hipSuccess && na > 1) le = hipEventRecord(pl->fork, ctx->stream);
        if (le == hipSuccess && na > 0)
            le = kind == EK_F64 ? class_launch<double>(pl, act[0], ctx->stream, false)
                                : class_launch<float>(pl, act[0], ctx->stream, false);
        for (int i = 1; i < na && le == hipSuccess; ++i) {
            hipStream_t st = pl->side[i - 1];
            le = hipStreamWaitEvent(st, pl->fork, 0);
            if (le == hipSuccess)
                le = kind == EK_F64 ? class_launch<double>(pl, act[i], st, false)
                                    : class_launch<float>(pl, act[i], st, false);
            if (le == hipSuccess) le = hipEventRecord(pl->join[i - 1], st);
        }
        for (int i = 1; i < na && le == hipSuccess; ++i) le = hipStreamWaitEvent(ctx->stream, pl->join[i - 1], 0);
        e = hipStreamEndCapture(ctx->stream, &pl->graph);
        if (e == hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        const int st = hip_check(ctx, e, "batched elimination plan");
        batched_elim_destroy(ctx, pl);
        return st;
    }
    *out = pl;
    return MODP_OK;
}

int batched_elim_run(modp_ctx *ctx, modp_elim_plan *pl)
{
    if (pl->count == 0) return MODP_OK;
    TimedLaunch tl(ctx, "batched_elim");
    return hip_check(ctx, hipGraphLaunch(pl->exec, ctx->stream), "hipGraphLaunch");
}

int batched_elim(modp_ctx *ctx, int op, int kind, const modp_elim_job *jobs, int64_t count)
{
    if (count <= 0) return MODP_OK;
    modp_elim_plan *pl = nullptr;
    MODP_TRY(batched_elim_plan(ctx, op, kind, jobs, count, &pl));
    int st = batched_elim_run(ctx, pl);
    // the plan owns the job table: finish before releasing it
    if (st == MODP_OK) st = hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    batched_elim_destroy(ctx, pl);
    return st;
}
