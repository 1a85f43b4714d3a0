// batched.hip -- thousands of independent small products / eliminations, each
// under its own prime, in one persistent launch fed by a device work queue.
//
// GEMM: the plan sorts jobs by cost, cuts each into output tiles whose shape
// the autotuner picks for the job's shape bucket, and captures
// [queue reset kernel, persistent kernel] in a hipGraph; every workgroup pulls
// tile indices from an atomic counter until the queue is empty.
//
// Elimination: one workgroup per instance (m, n <= 128) with the matrix held
// in LDS, the same pivot rule as the single-matrix path, one pass per pivot.
#include <algorithm>
#include <map>
#include <numeric>

#include "batched.h"
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
};

struct BItem {
    int job;
    int tile;  // tile index within the job: ti = tile / tiles_n, tj = tile % tiles_n
};

// Work items are grouped in classes (tile 64 or 32) x (reduction needed or
// not); each class is served by its own persistent kernel, which carries
// only that class's tile code (fewer registers, more resident waves), and
// pulls items from its own counter.
constexpr int NCLASS = 4;
struct modp_gemm_plan {
    int kind = EK_F64;
    int64_t count = 0, items = 0;
    int64_t cls_off[NCLASS + 1] = {0};  // item ranges of the classes
    BJob *djobs = nullptr;
    BItem *ditems = nullptr;
    int *dcounter = nullptr;  // one counter per class, 64 bytes apart
    int grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

__global__ void k_queue_reset(int *counter)
{
    if (threadIdx.x < NCLASS && blockIdx.x == 0) counter[threadIdx.x * 16] = 0;
}

template <typename T, int TILE, int RED>
__global__ __launch_bounds__(256, 2) void k_batched_gemm(const BJob *__restrict__ jobs,
                                                         const BItem *__restrict__ items, int64_t nitems,
                                                         int *__restrict__ counter)
{
    __shared__ F64Smem<TILE, TILE> sm;
    __shared__ int s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= nitems) break;
        const BItem item = items[it];
        const BJob &j = jobs[item.job];
        const int ti = item.tile / j.tiles_n, tj = item.tile % j.tiles_n;
        gemm_f64_tile<T, T, T, TILE, TILE, EPI_STORE, RED>(j.m, j.n, j.k, (const T *)j.A, j.lda, (const T *)j.B,
                                                           j.ldb, (T *)j.C, j.ldc, j.mp, 1, (int64_t)ti * TILE,
                                                           (int64_t)tj * TILE, sm);
    }
}

template <typename T>
void launch_class(int cls, const modp_gemm_plan *pl, hipStream_t st)
{
    const int64_t n = pl->cls_off[cls + 1] - pl->cls_off[cls];
    if (n <= 0) return;
    const BItem *it = pl->ditems + pl->cls_off[cls];
    int *ctr = pl->dcounter + cls * 16;
    // resident workgroups per CU by the kernels' register use: 3 for 64x64
    // tiles, 5 for 32x32
    const int64_t cap = (int64_t)pl->grid * (cls < 2 ? 3 : 5);
    const unsigned g = (unsigned)(n < cap ? n : cap);
    switch (cls) {
    case 0: k_batched_gemm<T, 64, RED_NONE><<<g, 256, 0, st>>>(pl->djobs, it, n, ctr); break;
    case 1: k_batched_gemm<T, 64, RED_STEP><<<g, 256, 0, st>>>(pl->djobs, it, n, ctr); break;
    case 2: k_batched_gemm<T, 32, RED_NONE><<<g, 256, 0, st>>>(pl->djobs, it, n, ctr); break;
    default: k_batched_gemm<T, 32, RED_STEP><<<g, 256, 0, st>>>(pl->djobs, it, n, ctr); break;
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
        d.tiles_n = (d.n + d.tile - 1) / d.tile;
    }
    // most expensive tiles first, within each class
    auto cost = [&](int64_t i) { return (int64_t)hj[i].tile * hj[i].tile * hj[i].k; };
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return cost(a) > cost(b); });
    auto cls_of = [&](const BJob &d) {
        const bool exact = (int64_t)d.R >= ((int64_t)d.k + 3) / 4;
        return (d.tile == 64 ? 0 : 2) + (exact ? 0 : 1);
    };
    std::vector<BItem> items;
    for (int c = 0; c < NCLASS; ++c) {
        pl->cls_off[c] = (int64_t)items.size();
        for (int64_t oi : order) {
            const BJob &d = hj[oi];
            if (d.m <= 0 || d.n <= 0 || cls_of(d) != c) continue;
            const int tiles = ((d.m + d.tile - 1) / d.tile) * d.tiles_n;
            for (int t = 0; t < tiles; ++t) items.push_back(BItem{(int)oi, t});
        }
    }
    pl->cls_off[NCLASS] = (int64_t)items.size();
    pl->items = (int64_t)items.size();
    hipError_t e = hipMalloc(&pl->djobs, sizeof(BJob) * (count > 0 ? count : 1));
    if (e == hipSuccess) e = hipMalloc(&pl->ditems, sizeof(BItem) * (items.empty() ? 1 : items.size()));
    if (e == hipSuccess) e = hipMalloc(&pl->dcounter, sizeof(int) * 64);
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
        k_queue_reset<<<1, 64, 0, ctx->stream>>>(pl->dcounter);
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
// batched elimination
// ---------------------------------------------------------------------------

namespace {

constexpr int EMAX = 128;
constexpr int EP = EMAX + 1;  // LDS row pitch (odd: column walks hit distinct banks)

struct EJob {
    const void *A; int64_t lda;
    void *U; int64_t ldu;
    void *L; int64_t ldl;
    int64_t *row_info;
    int64_t *col_info;
    int64_t *counts;
    int m, n;
    ModP mp;
};

struct ESmem {
    uint32_t S[EMAX * EP];
    uint32_t mult[EMAX];
    uint32_t pivval[EMAX];
    int pivcol[EMAX];
    int ipiv[EMAX];
    uint32_t wv[4];
    int wi[4];
    int flag;
    int item;
};

// argmax of S[r][col] for r in [row, m): largest value, lowest row; result
// in (*V, *I) for every thread.  Two barriers.
__device__ __forceinline__ void col_argmax(ESmem &s, int m, int row, int col, uint32_t *V, int *I)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t bv = 0;
    int bi = 0x7fffffff;
    for (int r = row + tid; r < m; r += 256) {
        const uint32_t x = s.S[r * EP + col];
        if (x > bv) { bv = x; bi = r; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t ov = __shfl_xor(bv, off);
        const int oi = __shfl_xor(bi, off);
        if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if (lane == 0) { s.wv[wave] = bv; s.wi[wave] = bi; }
    __syncthreads();
    uint32_t v = s.wv[0];
    int i = s.wi[0];
#pragma unroll
    for (int x = 1; x < 4; ++x)
        if (s.wv[x] > v || (s.wv[x] == v && s.wi[x] < i)) { v = s.wv[x]; i = s.wi[x]; }
    *V = v;
    *I = i;
    __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(256) void k_batched_elim(const EJob *__restrict__ jobs, int64_t njobs, int op,
                                                      int *__restrict__ counter)
{
    __shared__ ESmem s;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s.item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s.item;
        __syncthreads();
        if (it >= njobs) break;
        const EJob &j = jobs[it];
        const int m = j.m, n = j.n;
        const ModP mp = j.mp;
        const T *A = (const T *)j.A;
        for (int e = tid; e < m * n; e += 256) {
            const int r = e % m, c = e / m;
            s.S[r * EP + c] = to_res<T>(A[r + (int64_t)c * j.lda], mp);
        }
        __syncthreads();
        int row = 0, col = 0, last = n - 1, nr = 0, nc = 0;
        bool bad = false;
        while (row < m && col < n) {
            uint32_t V;
            int I;
            col_argmax(s, m, row, col, &V, &I);
            if (V == 0) {
                if (op != MODP_ELIM_PLUQ) { ++col; continue; }
                // pluq: whole remaining block zero?
                if (tid == 0) s.flag = 0;
                __syncthreads();
                const int rows = m - row, cols = n - col;
                for (int e = tid; e < rows * cols; e += 256)
                    if (s.S[(row + e % rows) * EP + col + e / rows]) s.flag = 1;
                __syncthreads();
                const int any = s.flag;
                __syncthreads();
                if (!any) {
                    // src tail: (c, n) for c = col..n-1, applied in order
                    for (int r = tid; r < m; r += 256)
                        for (int c = col; c < n - 1; ++c) {
                            uint32_t t = s.S[r * EP + c];
                            s.S[r * EP + c] = s.S[r * EP + n - 1];
                            s.S[r * EP + n - 1] = t;
                        }
                    if (tid == 0)
                        for (int c = col; c < n; ++c) {
                            j.col_info[2 * nc] = c + 1;
                            j.col_info[2 * nc + 1] = n;
                            ++nc;
                        }
                    __syncthreads();
                    break;
                }
                for (int r = tid; r < m; r += 256) {
                    uint32_t t = s.S[r * EP + col];
                    s.S[r * EP + col] = s.S[r * EP + last];
                    s.S[r * EP + last] = t;
                }
                if (tid == 0) {
                    j.col_info[2 * nc] = col + 1;
                    j.col_info[2 * nc + 1] = last + 1;
                }
                ++nc;
                --last;
                __syncthreads();
                continue;
            }
            const uint32_t inv = modinv(V, mp);
            if (inv == 0) { bad = true; break; }
            // swap rows row <-> I and normalise the pivot row from col on
            for (int c = tid; c < n; c += 256) {
                const uint32_t a = s.S[row * EP + c], b = s.S[I * EP + c];
                s.S[I * EP + c] = a;
                s.S[row * EP + c] = c >= col ? mulmod(b, inv, mp) : b;
            }
            if (tid == 0) {
                s.pivval[row] = V;
                s.pivcol[row] = col;
                s.ipiv[row] = I;
                if (op == MODP_ELIM_PLUQ && I != row) {
                    j.row_info[2 * nr] = row + 1;
                    j.row_info[2 * nr + 1] = I + 1;
                }
            }
            if (op == MODP_ELIM_PLUQ && I != row) ++nr;
            __syncthreads();
            if (op == MODP_ELIM_RREF) {
                // Gauss-Jordan: clear the pivot column in every other row
                for (int r = tid; r < m; r += 256) s.mult[r] = r == row ? 0u : s.S[r * EP + col];
                __syncthreads();
                const int cols = n - col;
                for (int e = tid; e < m * cols; e += 256) {
                    const int r = e / cols, c = col + e % cols;
                    const uint32_t l = s.mult[r];
                    if (l) s.S[r * EP + c] = submod(s.S[r * EP + c], mulmod(l, s.S[row * EP + c], mp), mp);
                }
            } else {
                const int rows = m - row - 1, cols = n - col - 1;
                for (int e = tid; e < rows * cols; e += 256) {
                    const int r = row + 1 + e / cols, c = col + 1 + e % cols;
                    const uint32_t l = s.S[r * EP + col];
                    if (l) s.S[r * EP + c] = submod(s.S[r * EP + c], mulmod(l, s.S[row * EP + c], mp), mp);
                }
            }
            __syncthreads();
            ++row;
            ++col;
        }
        const int rank = row;
        // outputs
        T *U = (T *)j.U;
        for (int e = tid; e < m * n; e += 256) {
            const int r = e % m, c = e / m;
            uint32_t v = 0;
            if (op == MODP_ELIM_RREF) v = r < rank ? s.S[r * EP + c] : 0u;
            else if (r < rank && c >= s.pivcol[r]) v = s.S[r * EP + c];
            U[r + (int64_t)c * j.ldu] = from_res<T>(v);
        }
        if (op != MODP_ELIM_RREF) {
            T *L = (T *)j.L;
            for (int e = tid; e < m * m; e += 256) {
                const int r = e % m, k = e / m;
                uint32_t v = 0;
                if (k < rank) {
                    if (r == k) v = s.pivval[k];
                    else if (r > k) v = s.S[r * EP + s.pivcol[k]];
                }
                L[r + (int64_t)k * j.ldl] = from_res<T>(v);
            }
        }
        if (tid == 0) {
            if (op == MODP_ELIM_LU) {
                for (int i = 0; i < m; ++i) j.row_info[i] = i + 1;
                for (int i = 0; i < rank; ++i) {
                    const int64_t t = j.row_info[i];
                    j.row_info[i] = j.row_info[s.ipiv[i]];
                    j.row_info[s.ipiv[i]] = t;
                }
                for (int i = 0; i < rank; ++i) j.col_info[i] = s.pivcol[i] + 1;
            } else if (op == MODP_ELIM_RREF) {
                for (int i = 0; i < rank; ++i) j.row_info[i] = s.pivcol[i] + 1;
            }
            j.counts[0] = op == MODP_ELIM_PLUQ ? nr : 0;
            j.counts[1] = op == MODP_ELIM_PLUQ ? nc : 0;
            j.counts[2] = bad ? -1 : rank;
        }
        __syncthreads();
    }
}

}  // namespace

int batched_elim(modp_ctx *ctx, int op, int kind, const modp_elim_job *jobs, int64_t count)
{
    if (count <= 0) return MODP_OK;
    if (kind == EK_I32) return set_error(ctx, MODP_ERR_DTYPE, "batched elimination: unsupported element type");
    std::vector<EJob> hj((size_t)count);
    std::vector<int64_t> order((size_t)count);
    std::iota(order.begin(), order.end(), 0);
    auto cost = [&](int64_t i) {
        const int64_t a = jobs[i].m, b = jobs[i].n;
        return a * b * (a < b ? a : b);
    };
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return cost(a) > cost(b); });
    for (int64_t i = 0; i < count; ++i) {
        const modp_elim_job &s = jobs[order[i]];
        EJob &d = hj[i];
        d.A = s.A; d.lda = s.lda; d.U = s.U; d.ldu = s.ldu; d.L = s.L; d.ldl = s.ldl;
        d.row_info = s.row_info; d.col_info = s.col_info; d.counts = s.counts;
        d.m = (int)s.m; d.n = (int)s.n;
        d.mp = make_modp(s.p);
    }
    int st = MODP_OK;
    char *buf = (char *)ws_get(ctx, WS_BATCH, sizeof(EJob) * count + 256, &st);
    if (!buf) return st;
    int *counter = (int *)buf;
    EJob *dj = (EJob *)(buf + 256);
    HIP_TRY(ctx, hipMemcpyAsync(dj, hj.data(), sizeof(EJob) * count, hipMemcpyHostToDevice, ctx->stream));
    k_queue_reset<<<1, 64, 0, ctx->stream>>>(counter);
    HIP_TRY(ctx, hipGetLastError());
    int64_t g = (int64_t)ctx->num_cus * 2;
    if (g > count) g = count;
    {
        TimedLaunch tl(ctx, "batched_elim");
        if (kind == EK_F64)
            k_batched_elim<double><<<(unsigned)g, 256, 0, ctx->stream>>>(dj, count, op, counter);
        else
            k_batched_elim<float><<<(unsigned)g, 256, 0, ctx->stream>>>(dj, count, op, counter);
        HIP_TRY(ctx, hipGetLastError());
    }
    // the job table lives in a reused workspace: finish before returning
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MODP_OK;
}
