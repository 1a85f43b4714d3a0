// common.h -- shared device helpers and internal host interfaces of
// libmodp_hip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <map>
#include <string>
#include <vector>

#include "../../include/modp.h"

// ---------------------------------------------------------------------------
// modulus parameters passed by value to every kernel
// ---------------------------------------------------------------------------
struct ModP {
    int64_t p;      // modulus, 2 <= p <= 2^26
    double pd;      // (double)p
    double pinv;    // 1/p rounded to double
    float pinvf;    // 1/p rounded to float
    uint32_t p32;   // p as u32
    int small;      // p < 2^16: products fit in u32
    int half;       // floor(p/2): balanced representative bound
};

static inline ModP make_modp(int64_t p)
{
    ModP m;
    m.p = p;
    m.pd = (double)p;
    m.pinv = 1.0 / (double)p;
    m.pinvf = 1.0f / (float)p;
    m.p32 = (uint32_t)p;
    m.small = p < 65536 ? 1 : 0;
    m.half = (int)(p / 2);
    return m;
}

// x in (-2^53, 2^53), integral -> x mod p in [0,p).  q = rint(x/p) is off
// by at most one (fl(x * pinv) is within 2^53 * 2^-52 / p <= 2/3 of x/p), so
// r = x - q p is exact (|r| < 2^53) and |r| < 1.5p + 1; three fix-ups bring it
// into [0,p).  (Round-to-nearest runs at twice the rate of floor here.)
__device__ __forceinline__ double dmod(double x, const ModP &mp)
{
    double q = rint(x * mp.pinv);
    double r = fma(-q, mp.pd, x);
    r = r < 0.0 ? r + mp.pd : r;
    r = r >= mp.pd ? r - mp.pd : r;
    r = r < 0.0 ? r + mp.pd : r;
    return r;
}

// a, b in [0,p) -> a*b mod p
__device__ __forceinline__ uint32_t mulmod(uint32_t a, uint32_t b, const ModP &mp)
{
    if (mp.small) {
        uint32_t x = a * b;
        uint32_t q = (uint32_t)((float)x * mp.pinvf);
        int32_t r = (int32_t)(x - q * mp.p32);
        r += r < 0 ? (int32_t)mp.p32 : 0;
        r -= r >= (int32_t)mp.p32 ? (int32_t)mp.p32 : 0;
        return (uint32_t)r;
    }
    double x = (double)a * (double)b;
    return (uint32_t)dmod(x, mp);
}

__device__ __forceinline__ uint32_t addmod(uint32_t a, uint32_t b, const ModP &mp)
{
    uint32_t s = a + b;
    return s >= mp.p32 ? s - mp.p32 : s;
}

__device__ __forceinline__ uint32_t submod(uint32_t a, uint32_t b, const ModP &mp)
{
    return a >= b ? a - b : a + mp.p32 - b;
}

// a^{-1} mod p by the extended Euclidean algorithm
// (src/CuModMatrix/rref_lu_pluq/pluq_kernels.jl:11-31).  Returns 0 when
// gcd(a,p) != 1, which the callers report as "not invertible".
// The Bezout coefficients stay within [-p, p], so int32 holds every
// intermediate for p <= 2^26; the quotient comes from a double division with
// one correction step.
__device__ __forceinline__ uint32_t modinv(uint32_t a, const ModP &mp)
{
    int32_t inv = 0, ninv = 1, rem = (int32_t)mp.p32, nrem = (int32_t)a;
    while (nrem != 0) {
        int32_t q = (int32_t)((double)rem / (double)nrem);
        int32_t r = rem - q * nrem;
        if (r < 0) { q -= 1; r += nrem; }
        else if (r >= nrem) { q += 1; r -= nrem; }
        int32_t t = inv - q * ninv; inv = ninv; ninv = t;
        rem = nrem; nrem = r;
    }
    if (rem != 1) return 0;
    if (inv < 0) inv += (int32_t)mp.p32;
    return (uint32_t)inv;
}

// XCD-aware tile order for a 1-D grid of 8 * ceil(tiles/8) workgroups.
// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8), each with its own L2: give XCD x the contiguous range
// [x*per, (x+1)*per) of the tile sequence, and walk that sequence in groups
// of GROUP tile-rows so concurrent tiles of one XCD share A and B panels.
// Returns false for the padding workgroups.
__device__ __forceinline__ bool tile_of_block(int64_t tiles_m, int64_t tiles_n, int64_t *tm, int64_t *tn)
{
    constexpr int NXCD = 8, GROUP = 8;
    const int64_t total = tiles_m * tiles_n;
    const int64_t per = (total + NXCD - 1) / NXCD;
    const int64_t b = blockIdx.x;
    const int64_t t = (b % NXCD) * per + b / NXCD;
    if (t >= total) return false;
    const int64_t gsize = (int64_t)GROUP * tiles_n;
    const int64_t g = t / gsize, first = g * GROUP;
    const int64_t gm = tiles_m - first < GROUP ? tiles_m - first : GROUP;
    const int64_t r = t % gsize;
    *tm = first + r % gm;
    *tn = r / gm;
    return true;
}

static inline unsigned xcd_grid(int64_t tiles) { return (unsigned)(((tiles + 7) / 8) * 8); }

// a^{-1} mod p for p < 2^24: extended Euclid whose quotients come from a
// float reciprocal (exact operands below 2^24, estimate corrected to the
// exact quotient); larger p use the double-division variant.
__device__ __forceinline__ uint32_t modinv_fast(uint32_t a, const ModP &mp)
{
    if (mp.p >= (1 << 24)) return modinv(a, mp);
    int32_t inv = 0, ninv = 1, rem = (int32_t)mp.p32, nrem = (int32_t)a;
    while (nrem != 0) {
        int32_t q = (int32_t)((float)rem * __builtin_amdgcn_rcpf((float)nrem));
        int32_t r = rem - q * nrem;
        while (r < 0) { --q; r += nrem; }
        while (r >= nrem) { ++q; r -= nrem; }
        int32_t t = inv - q * ninv; inv = ninv; ninv = t;
        rem = nrem; nrem = r;
    }
    if (rem != 1) return 0;
    if (inv < 0) inv += (int32_t)mp.p32;
    return (uint32_t)inv;
}

// ---------------------------------------------------------------------------
// wavefront reductions by DPP
// ---------------------------------------------------------------------------

template <int CTRL> __device__ __forceinline__ uint32_t dpp32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, false);
}
template <int CTRL> __device__ __forceinline__ uint64_t dpp64(uint64_t v)
{
    return ((uint64_t)dpp32<CTRL>((uint32_t)(v >> 32)) << 32) | dpp32<CTRL>((uint32_t)v);
}
template <typename T> __device__ __forceinline__ T kmax(T a, T b) { return a > b ? a : b; }

// wave64 maximum: DPP within each row of 16 lanes (quad swaps, half-row and
// row mirrors), then the four row results through readlane
__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
    v = kmax(v, dpp32<0xB1>(v));   // quad_perm [1,0,3,2]
    v = kmax(v, dpp32<0x4E>(v));   // quad_perm [2,3,0,1]
    v = kmax(v, dpp32<0x141>(v));  // row_half_mirror
    v = kmax(v, dpp32<0x140>(v));  // row_mirror
    const uint32_t r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const uint32_t r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return kmax(kmax(r0, r1), kmax(r2, r3));
}
__device__ __forceinline__ uint64_t wave_max(uint64_t v)
{
    v = kmax(v, dpp64<0xB1>(v));
    v = kmax(v, dpp64<0x4E>(v));
    v = kmax(v, dpp64<0x141>(v));
    v = kmax(v, dpp64<0x140>(v));
    // readlane returns a signed int: both halves go through uint32_t so a low
    // word with bit 31 set is not sign-extended into the high word
    uint64_t r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        r[i] = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((uint32_t)(v >> 32), 16 * i) << 32) |
               (uint64_t)(uint32_t)__builtin_amdgcn_readlane((uint32_t)v, 16 * i);
    return kmax(kmax(r[0], r[1]), kmax(r[2], r[3]));
}

// storage element -> residue in [0,p)
template <typename T> __device__ __forceinline__ uint32_t to_res(T x, const ModP &mp);
template <> __device__ __forceinline__ uint32_t to_res<double>(double x, const ModP &mp)
{
    return (uint32_t)dmod(x, mp);
}
template <> __device__ __forceinline__ uint32_t to_res<float>(float x, const ModP &mp)
{
    return (uint32_t)dmod((double)x, mp);
}
template <> __device__ __forceinline__ uint32_t to_res<int32_t>(int32_t x, const ModP &mp)
{
    // working matrices hold residues already
    (void)mp;
    return (uint32_t)x;
}

template <typename T> __device__ __forceinline__ T from_res(uint32_t r);
template <> __device__ __forceinline__ double from_res<double>(uint32_t r) { return (double)r; }
template <> __device__ __forceinline__ float from_res<float>(uint32_t r) { return (float)r; }
template <> __device__ __forceinline__ int32_t from_res<int32_t>(uint32_t r) { return (int32_t)r; }

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct KernelTiming {
    std::string name;
    hipEvent_t a, b;
};

// GEMM launch configuration chosen by the heuristics or the autotuner
struct GemmConfig {
    int form = 0;     // MODP_FORM_PLAIN_F64 or MODP_FORM_LIMB_I8
    int tile_m = 64;  // output tile: 64x64, 128x128, or (limb form) 128x64
    int tile_n = 64;
    int ksplit = 1;   // K slices run concurrently and summed mod p
};

// autotuner cache entries: key = (element kinds, bits of p, log2 buckets of
// m, n, k) for single GEMMs; (kind, log2 buckets of count and mean m, n, k)
// with kind + 16 for batched plans
struct TuneKey {
    int v[7];
    bool operator<(const TuneKey &o) const
    {
        for (int i = 0; i < 7; ++i)
            if (v[i] != o.v[i]) return v[i] < o.v[i];
        return false;
    }
};
struct TuneEntry {
    GemmConfig cfg;     // single GEMM
    double thr = 0.0;   // batched plan: 64-tile threshold
    double ms = 0.0;    // measured time of the winner
};

constexpr int WS_COUNT = 21;

struct modp_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    // grow-only device workspaces
    void *ws[WS_COUNT] = {nullptr};
    size_t ws_size[WS_COUNT] = {0};
    // pinned host staging for small device->host readbacks
    void *pinned = nullptr;
    size_t pinned_size = 0;
    // kernel timing
    bool timing = false;
    std::vector<KernelTiming> timings;
    std::vector<hipEvent_t> event_pool;
    hipEvent_t slots[16] = {nullptr};
    // table of modular inverses for the current elimination modulus
    uint32_t *invtab = nullptr;
    int64_t invtab_p = 0;
    size_t invtab_cap = 0;
    // autotuner: MODP_TUNE_OFF / _CACHED / _ON_MISS and its cache
    int tune_mode = 0;
    std::map<TuneKey, TuneEntry> tune_cache;
    // side streams (default priority) and fork/join events for independent
    // small launches; created on first use
    static constexpr int NSIDE = 8;
    hipStream_t side[NSIDE] = {nullptr};
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[NSIDE] = {nullptr};
    // device address of the library's static scratch array on this
    // context's device (modp_gemm_noalloc), and a flag set while such a call
    // runs (no timing events are taken then)
    int8_t *noalloc_scratch = nullptr;
    bool in_noalloc = false;
};

// workspace slots
enum {
    WS_ALIMB = 0,
    WS_BLIMB = 1,
    WS_WORK = 2,
    WS_PIV = 3,
    WS_UBUF = 4,
    WS_MISC = 5,
    WS_BATCH = 6,
    WS_SCRATCH = 7,
    WS_SLICES = 8,   // K-slice partial results
    WS_TUNE_A = 9,   // synthetic operands of modp_autotune_gemm
    WS_TUNE_B = 10,
    WS_TUNE_C = 11,
    WS_TRI_FLAG = 12,  // triangular inverse: zero-diagonal flag
    WS_TRI_TMP = 13,   //   GEMM temporary
    WS_TRI_T = 14,     //   int32 copy of the triangle
    WS_TRI_X = 15,     //   int32 inverse
    WS_TRI_X2 = 16,    //   second inverse (matrix inverse: L^{-1})
    WS_INV_U = 17,     // matrix inverse: PLUQ factors
    WS_INV_L = 18,
    WS_PERM = 19,      // lines moved by a permutation
    WS_GPOS = 20       // row positions of the W-resident elimination panel
};

// error plumbing: every internal function returns a modp_status
#define MODP_TRY(expr)                    \
    do {                                  \
        int _st = (expr);                 \
        if (_st != MODP_OK) return _st;   \
    } while (0)

int set_error(modp_ctx *ctx, int code, const char *fmt, ...);
int hip_check(modp_ctx *ctx, hipError_t e, const char *what);
#define HIP_TRY(ctx, expr) MODP_TRY(hip_check((ctx), (expr), #expr))

void *ws_get(modp_ctx *ctx, int slot, size_t bytes, int *status);
void *pinned_get(modp_ctx *ctx, size_t bytes, int *status);

// timing hooks around hot kernel launches
struct TimedLaunch {
    modp_ctx *ctx;
    KernelTiming *t;
    TimedLaunch(modp_ctx *c, const char *name);
    ~TimedLaunch();
};

// ---------------------------------------------------------------------------
// internal kernel launchers (defined in the .hip files)
// ---------------------------------------------------------------------------

// element kind of a matrix operand
enum ElemKind { EK_F64 = 0, EK_F32 = 1, EK_I32 = 2 };

// epilogue of the GEMM cores
enum GemmEpi {
    EPI_STORE = 0,  // C = A*B mod p
    EPI_SUB = 1,    // C = (C - A*B) mod p
    EPI_ADD = 2     // C = (C + A*B) mod p
};

struct GemmArgs {
    int64_t m, n, k;
    const void *A; int64_t lda; int ka;   // ka: ElemKind of A
    const void *B; int64_t ldb; int kb;
    void *C; int64_t ldc; int kc;
    int epi;
    ModP mp;
    // optional device flag: when nonzero at launch time the kernels return
    // without touching memory (the elimination's optimistic pipeline)
    const int *skip = nullptr;
};

// heuristic configurations (gemm.hip)
int gemm_plain_f64(modp_ctx *ctx, const GemmArgs &g);
int gemm_limb_i8(modp_ctx *ctx, const GemmArgs &g);
int gemm_rns_i8(modp_ctx *ctx, const GemmArgs &g);
// the limb form with its digit planes in a caller's workspace (no allocation,
// one K slice); gemm_limb_scratch_bytes is the workspace that product needs
struct LimbScratch {
    int8_t *base;
    size_t bytes;
};
size_t gemm_limb_scratch_bytes(const modp_ctx *ctx, const GemmArgs &g);
// the multi-modular (RNS) int8 form (gemm_rns.hip): moduli needed for an
// exact k-deep product mod p (0: none suffice), and the product with a
// tile x tile output tile
int rns_count(int64_t p, int64_t k);
// residue planes of the RNS form: 1 for 2 <= p <= 255 (one int8 product
// modulo p itself), else rns_count
int rns_planes(int64_t p, int64_t k);
int gemm_rns_run(modp_ctx *ctx, const GemmArgs &g, int tile);
constexpr int64_t RNS_KCHUNK = 130944;
int gemm_limb_i8_in(modp_ctx *ctx, const GemmArgs &g, const LimbScratch &ws);
int limb_count(int64_t p);
int choose_form(int64_t m, int64_t n, int64_t k, int64_t p);
GemmConfig gemm_default_config(const modp_ctx *ctx, const GemmArgs &g, int form);
bool gemm_config_valid(const GemmArgs &g, const GemmConfig &cfg);
int gemm_run(modp_ctx *ctx, const GemmArgs &g, const GemmConfig &cfg);
int gemv_run(modp_ctx *ctx, int kind, int64_t m, int64_t k, const void *A, int64_t lda, const void *x, void *y,
             int64_t p);

// autotuner (tune.hip)
int gemm_auto(modp_ctx *ctx, const GemmArgs &g);
int tune_gemm(modp_ctx *ctx, const GemmArgs &g, GemmConfig *best, double *best_ms);
TuneKey tune_key_gemm(const GemmArgs &g);
GemmConfig gemm_lookup_config(const modp_ctx *ctx, const GemmArgs &g, int *tuned);
int tune_export(modp_ctx *ctx, char *buf, size_t cap, size_t *len);

// device peak microbenchmarks (peak.hip)
int probe_peak(modp_ctx *ctx, int kind, double *rate);
int tune_import(modp_ctx *ctx, const char *text);

int util_mod(modp_ctx *ctx, int kind, int64_t rows, int64_t cols, void *X, int64_t ldx, int64_t N);
int util_rand(modp_ctx *ctx, int kind, int64_t rows, int64_t cols, void *X, int64_t ldx, int64_t p,
              uint64_t seed, uint64_t offset);
int util_eye(modp_ctx *ctx, int kind, int64_t rows, int64_t cols, void *X, int64_t ldx);
int util_convert(modp_ctx *ctx, int kin, int64_t rows, int64_t cols, const void *X, int64_t ldx,
                 int kout, void *Y, int64_t ldy, const ModP &mp);
int util_perm(modp_ctx *ctx, int kind, int rowwise, int64_t rows, int64_t cols, void *X, int64_t ldx,
              const int64_t *stack_host, int64_t nswaps, int inverse);
int util_crt(modp_ctx *ctx, int kind, int64_t rows, int64_t cols, int r, const void *const *X, const int64_t *ldx,
             const int64_t *p, int sign, uint32_t *Y, int64_t ldy, int64_t words);
int util_crt2(modp_ctx *ctx, int kind, int64_t rows, int64_t cols, const void *X1, int64_t ld1, int64_t p1,
              const void *X2, int64_t ld2, int64_t p2, void *Y, int64_t ldy);

// elimination driver (elim.hip)
enum ElimMode { ELIM_PLUQ = 0, ELIM_LU = 1, ELIM_RREF = 2 };
struct ElimResult {
    std::vector<int64_t> row_swaps, col_swaps, perm, pivcols;
    int64_t rank = 0;
};
int elim_run(modp_ctx *ctx, int mode, int form, int kind, int64_t m, int64_t n,
             const void *A, int64_t lda, void *U, int64_t ldu, void *L, int64_t ldl,
             int64_t p, ElimResult &res);
int elim_rref(modp_ctx *ctx, int form, int kind, int64_t m, int64_t n, const void *A, int64_t lda, void *R,
              int64_t ldr, int64_t p, ElimResult &res);

// inverses (trinv.hip)
int tri_inverse(modp_ctx *ctx, int kind, int upper, int64_t m, int64_t n, const void *A, int64_t lda, void *Ainv,
                int64_t ldai, int64_t p);
int mat_inverse(modp_ctx *ctx, int form, int kind, int64_t n, const void *A, int64_t lda, void *Ainv, int64_t ldai,
                int64_t p, int *invertible);
