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
};

struct modp_elim_plan {
    int op = 0, kind = EK_F64;
    int64_t count = 0;
    EJob *djobs = nullptr;
    int *dcounter = nullptr;
    size_t slot = 0;  // LDS bytes per workgroup (one wavefront)
    int bk_off = 0;   // offset of the pivot bookkeeping in the slot
    int grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

constexpr int EMAX = 128;
constexpr int EBOOK = EMAX * (4 + 4 + 2 + 2);  // pivot values and inverses, pivot columns, row swaps

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
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
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
// congruent; q = trunc(fl(t) fl(1/p)) is within one of floor(t/p) (the
// relative error of the float quotient is below 3 2^-24 and t/p <= 2^16), so
// r = t - q p lies in [-p, 2p) and two fix-ups finish it.  All products are
// 24-bit integer multiplies.  Otherwise c = pr and the FP64 reduction runs
// (|x - l pr| < 2^52 is exact).
template <bool SMALL>
__device__ __forceinline__ uint32_t row_update(uint32_t x, uint32_t l, uint32_t c, const ModP &mp)
{
    if (SMALL) {
        const uint32_t t = __umul24(l, c) + x;
        const uint32_t q = (uint32_t)((float)t * mp.pinvf);
        int32_t r = (int32_t)(t - __umul24(q, mp.p32));
        r += r < 0 ? (int32_t)mp.p32 : 0;
        r -= r >= (int32_t)mp.p32 ? (int32_t)mp.p32 : 0;
        return (uint32_t)r;
    }
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
    // a < 2^20: fl(a) is exact and the float quotient is within one
    static __device__ __forceinline__ uint32_t reduce(S a, const ModP &mp)
    {
        const uint32_t q = (uint32_t)((float)a * mp.pinvf);
        int32_t r = (int32_t)(a - __umul24(q, mp.p32));
        r += r < 0 ? (int32_t)mp.p32 : 0;
        r -= r >= (int32_t)mp.p32 ? (int32_t)mp.p32 : 0;
        return (uint32_t)r;
    }
};
template <> struct Acc<1> {  // FP64 sums, exact below 2^53
    typedef double V;        // stored: v, subtracted
    static __device__ __forceinline__ V from(uint32_t v, const ModP &) { return (double)v; }
    typedef double S;
    static __device__ __forceinline__ S init(uint32_t x) { return (double)x; }
    static __device__ __forceinline__ S madd(S a, uint32_t l, V v) { return fma(-(double)l, v, a); }
    static __device__ __forceinline__ uint32_t reduce(S a, const ModP &mp) { return (uint32_t)dmod(a, mp); }
};

// Trailing columns pe..n-1 of the pending pivots (rows prow..prow+np-1) on
// lanes over columns: lane q owns columns pe+q and (WIDE) pe+64+q; lanes
// past the last column repeat it, computing and storing the same value.
template <typename S, int KB, int ACC, bool WIDE>
__device__ __forceinline__ void flush_cols(S *M, int pitch, int m, int n, int op, int prow, int np, int pe,
                                           const int *pc, const uint32_t *inv, const uint32_t *lq0,
                                           const uint32_t *lq1, const ModP &mp)
{
    typedef Acc<ACC> A;
    constexpr bool SMALL = sizeof(S) < 4;
    const int lane = threadIdx.x;
    const int c0 = min(pe + lane, n - 1), c1 = min(pe + 64 + lane, n - 1);
    typename A::V v0[KB], v1[KB];
    // pivot rows, in order
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
    // every other row (rref: the panel's pivot rows too, by the later pivots
    // only); four rows' loads are issued before their stores
    const bool rref = op == MODP_ELIM_RREF;
    constexpr int RB = 4;
    for (int r0 = rref ? 0 : prow + np; r0 < m; r0 += RB) {
        uint32_t x0[RB], x1[RB];
#pragma unroll
        for (int t = 0; t < RB; ++t) {
            const int r = r0 + t;
            x0[t] = x1[t] = 0u;
            if (r < m) {
                x0[t] = M[r * pitch + c0];
                if (WIDE) x1[t] = M[r * pitch + c1];
            }
        }
#pragma unroll
        for (int t = 0; t < RB; ++t) {
            const int r = r0 + t;
            if (r < m) {
                typename A::S a0 = A::init(x0[t]), a1 = A::init(x1[t]);
                const int own = r - prow;  // rref: the panel's pivot row own (0 <= own < np)
#pragma unroll
                for (int jj = 0; jj < KB; ++jj) {
                    if (jj < np) {
                        uint32_t l = (uint32_t)__builtin_amdgcn_readlane((int)(r < 64 ? lq0[jj] : lq1[jj]), r & 63);
                        l = (own >= 0 && own < np && jj <= own) ? 0u : l;
                        a0 = A::madd(a0, l, v0[jj]);
                        if (WIDE) a1 = A::madd(a1, l, v1[jj]);
                    }
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
            flush_cols<S, KB, ACC, true>(M, pitch, m, n, op, prow, np, pe, pc, inv, lq0, lq1, mp);
        else
            flush_cols<S, KB, ACC, false>(M, pitch, m, n, op, prow, np, pe, pc, inv, lq0, lq1, mp);
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
    const int lane = threadIdx.x;
    const int m = j.m, n = j.n, pitch = j.pitch;
    const ModP &mp = j.mp;
    const T *A = (const T *)j.A;
    // column-major input, lanes over rows (two rows per lane at most):
    // 16 columns of loads in flight before their LDS stores
    for (int c0 = 0; c0 < n; c0 += 16) {
        T x0[16], x1[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const bool cv = c0 + t < n;
            x0[t] = (cv && lane < m) ? A[lane + (int64_t)(c0 + t) * j.lda] : T(0);
            x1[t] = (cv && lane + 64 < m) ? A[lane + 64 + (int64_t)(c0 + t) * j.lda] : T(0);
        }
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            if (c0 + t < n && lane < m) M[lane * pitch + c0 + t] = (S)to_res<T>(x0[t], mp);
            if (c0 + t < n && lane + 64 < m) M[(lane + 64) * pitch + c0 + t] = (S)to_res<T>(x1[t], mp);
        }
    }
    wave_sync();
    int row = 0, col = 0, last = n - 1, nr = 0, nc = 0;
    int prow = 0;               // first pivot row whose trailing update is pending
    int pe = KB < n ? KB : n;   // end of the current panel's columns
    bool bad = false, ended = false;
    while (row < m && col < n) {
        // largest residue of column col in rows row..m-1, lowest row on ties
        key_t mine = 0;
        if (lane >= row && lane < m) mine = K::make(M[lane * pitch + col], lane);
        if (lane + 64 >= row && lane + 64 < m) {
            const key_t k = K::make(M[(lane + 64) * pitch + col], lane + 64);
            mine = k > mine ? k : mine;
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
            const uint32_t inv = modinv_fast(bv, mp);
            if (inv == 0) { bad = true; break; }
            const int I = K::row(best);
            // swap rows row <-> I (every column) and normalise the pivot row
            // over the panel (the flush normalises the trailing columns)
            for (int c = lane; c < n; c += 64) {
                const uint32_t a = M[row * pitch + c], b = M[I * pitch + c];
                M[I * pitch + c] = (S)a;
                M[row * pitch + c] = (S)((c >= col && c < pe) ? mulmod_c<SMALL>(b, inv, mp) : b);
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
                uint32_t kp[KB];
#pragma unroll
                for (int t = 0; t < KB; ++t) {
                    const int c = col + 1 + t;
                    uint32_t v = c < pe ? (uint32_t)M[row * pitch + c] : 0u;
                    kp[t] = SMALL ? (v ? mp.p32 - v : 0u) : v;
                }
                const int i0 = op == MODP_ELIM_RREF ? 0 : row + 1;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int r = lane + 64 * h;
                    if (r < m && r >= i0 && r != row) {
                        const uint32_t l = M[r * pitch + col];
                        if (l) {
#pragma unroll
                            for (int t = 0; t < KB; ++t) {
                                const int c = col + 1 + t;
                                if (c < pe) M[r * pitch + c] = (S)row_update<SMALL>(M[r * pitch + c], l, kp[t], mp);
                            }
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
    // outputs, column-major: lanes over rows write consecutive elements
    T *U = (T *)j.U;
    for (int c = 0; c < n; ++c)
        for (int r = lane; r < m; r += 64) {
            uint32_t v = 0;
            if (op == MODP_ELIM_RREF) v = r < rank ? (uint32_t)M[r * pitch + c] : 0u;
            else if (r < rank && c >= pivcol[r]) v = M[r * pitch + c];
            U[r + (int64_t)c * j.ldu] = from_res<T>(v);
        }
    if (op != MODP_ELIM_RREF) {
        T *L = (T *)j.L;
        for (int k = 0; k < m; ++k) {
            const int pc = k < rank ? pivcol[k] : 0;
            const uint32_t pv = k < rank ? pivval[k] : 0u;
            for (int r = lane; r < m; r += 64) {
                uint32_t v = 0;
                if (k < rank) v = r == k ? pv : (r > k ? (uint32_t)M[r * pitch + pc] : 0u);
                L[r + (int64_t)k * j.ldl] = from_res<T>(v);
            }
        }
    }
    if (lane == 0) {
        if (op == MODP_ELIM_LU) {
            for (int q = 0; q < m; ++q) j.row_info[q] = q + 1;
            for (int q = 0; q < rank; ++q) {
                const int64_t t = j.row_info[q];
                j.row_info[q] = j.row_info[ipiv[q]];
                j.row_info[ipiv[q]] = t;
            }
            for (int q = 0; q < rank; ++q) j.col_info[q] = pivcol[q] + 1;
        } else if (op == MODP_ELIM_RREF) {
            for (int q = 0; q < rank; ++q) j.row_info[q] = pivcol[q] + 1;
        }
        j.counts[0] = op == MODP_ELIM_PLUQ ? nr : 0;
        j.counts[1] = op == MODP_ELIM_PLUQ ? nc : 0;
        j.counts[2] = bad ? -1 : rank;
    }
    wave_sync();
}

// Persistent kernel: each single-wavefront workgroup takes instances from
// one device queue (the job table is sorted by cost, largest first), the
// next index fetched while the current instance runs.
template <typename T>
__global__ __launch_bounds__(64) void k_elim_wave(const EJob *__restrict__ jobs, int64_t njobs, int op,
                                                  int *__restrict__ counter, int bk_off)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char esm[];
    int it = 0;
    if (threadIdx.x == 0) it = atomicAdd(counter, 1);
    it = __builtin_amdgcn_readfirstlane(it);
    while (it < njobs) {
        int next = 0;
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        const EJob j = jobs[it];
        if (j.cls == 0) elim_instance<T, uint8_t, 8, 0>(j, op, esm, bk_off);
        else if (j.cls == 1) elim_instance<T, uint16_t, 8, 1>(j, op, esm, bk_off);
        else elim_instance<T, uint32_t, 1, 1>(j, op, esm, bk_off);
        it = __builtin_amdgcn_readfirstlane(next);
    }
}

template <typename T> hipError_t allow_elim_slot(size_t slot)
{
    // the dynamic LDS can exceed the default 64 KiB (32-bit residues of
    // 128 x 128 instances)
    return hipFuncSetAttribute((const void *)k_elim_wave<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)slot);
}

template <typename T> hipError_t launch_elim(const modp_elim_plan *pl, hipStream_t st)
{
    k_elim_wave<T><<<(unsigned)pl->grid, 64, pl->slot, st>>>(pl->djobs, pl->count, pl->op, pl->dcounter, pl->bk_off);
    return hipGetLastError();
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
    if (pl->djobs) (void)hipFree(pl->djobs);
    if (pl->dcounter) (void)hipFree(pl->dcounter);
    delete pl;
    return MODP_OK;
}

// The plan sorts the instances by cost (largest first, so the queue ends
// with short ones), gives each its LDS element and row pitch, sizes the LDS
// slot for the largest, uploads the job table once and captures
// [queue reset, one persistent kernel] in a hipGraph.
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
    // cost: multiply-adds of the updates, the 32-bit class at twice the rate
    auto cost = [&](int64_t i) {
        const int64_t a = jobs[i].m, b = jobs[i].n;
        return a * b * (a < b ? a : b) * (elim_class(jobs[i].p) == 2 ? 2 : 1);
    };
    std::vector<int64_t> order((size_t)(count > 0 ? count : 0));
    std::iota(order.begin(), order.end(), (int64_t)0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return cost(a) > cost(b); });
    size_t mat_bytes = 16;
    for (int64_t oi : order) {
        const modp_elim_job &s = jobs[oi];
        EJob d;
        d.A = s.A; d.lda = s.lda; d.U = s.U; d.ldu = s.ldu; d.L = s.L; d.ldl = s.ldl;
        d.row_info = s.row_info; d.col_info = s.col_info; d.counts = s.counts;
        d.m = (int)s.m; d.n = (int)s.n;
        d.cls = elim_class(s.p);
        const int eb = d.cls == 0 ? 1 : (d.cls == 1 ? 2 : 4);
        d.pitch = elim_pitch(d.n, eb);
        d.mp = make_modp(s.p);
        mat_bytes = std::max(mat_bytes, (size_t)d.m * (size_t)d.pitch * (size_t)eb);
        hj.push_back(d);
    }
    pl->bk_off = (int)((mat_bytes + 15) & ~(size_t)15);
    pl->slot = (size_t)pl->bk_off + EBOOK;
    // persistent grid: as many single-wavefront workgroups as the LDS holds
    const size_t lds_cu = 160 * 1024;
    int64_t per_cu = (int64_t)(lds_cu / pl->slot);
    if (per_cu > 16) per_cu = 16;
    if (per_cu < 1) per_cu = 1;
    int64_t g = per_cu * ctx->num_cus;
    if (g > count) g = count;
    pl->grid = (int)(g > 0 ? g : 1);
    hipError_t e = hipMalloc(&pl->djobs, sizeof(EJob) * (count > 0 ? count : 1));
    if (e == hipSuccess) e = hipMalloc(&pl->dcounter, 256);
    if (e == hipSuccess && count > 0)
        e = hipMemcpyAsync(pl->djobs, hj.data(), sizeof(EJob) * count, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // hj is released on return
    if (e == hipSuccess) e = kind == EK_F64 ? allow_elim_slot<double>(pl->slot) : allow_elim_slot<float>(pl->slot);
    if (e == hipSuccess) e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        k_elim_counter_reset<<<1, 64, 0, ctx->stream>>>(pl->dcounter);
        const hipError_t le = kind == EK_F64 ? launch_elim<double>(pl, ctx->stream) : launch_elim<float>(pl, ctx->stream);
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
