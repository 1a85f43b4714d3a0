// api.hip -- the C ABI of libmodp_hip.so: context, memory, argument checks
// (mirroring the reference's exceptions as status codes) and dispatch.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "common.h"
#include "batched.h"

#define MODP_VERSION "modp_hip 0.1.0 (gfx950)"

// ---------------------------------------------------------------------------
// errors, workspaces, timing
// ---------------------------------------------------------------------------

int set_error(modp_ctx *ctx, int code, const char *fmt, ...)
{
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->last_error = buf;
    }
    return code;
}

int hip_check(modp_ctx *ctx, hipError_t e, const char *what)
{
    if (e == hipSuccess) return MODP_OK;
    if (e == hipErrorOutOfMemory) return set_error(ctx, MODP_ERR_OOM, "%s: out of device memory", what);
    return set_error(ctx, MODP_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

void *ws_get(modp_ctx *ctx, int slot, size_t bytes, int *status)
{
    if (bytes == 0) bytes = 256;
    if (ctx->ws_size[slot] >= bytes) return ctx->ws[slot];
    if (ctx->ws[slot]) {
        // in-flight kernels may still read the old buffer
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(ctx->ws[slot]);
        ctx->ws[slot] = nullptr;
        ctx->ws_size[slot] = 0;
    }
    size_t sz = (bytes + 4095) & ~(size_t)4095;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
        *status = hip_check(ctx, e, "workspace allocation");
        return nullptr;
    }
    ctx->ws[slot] = p;
    ctx->ws_size[slot] = sz;
    return p;
}

void *pinned_get(modp_ctx *ctx, size_t bytes, int *status)
{
    if (ctx->pinned_size >= bytes) return ctx->pinned;
    if (ctx->pinned) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipHostFree(ctx->pinned);
        ctx->pinned = nullptr;
        ctx->pinned_size = 0;
    }
    size_t sz = bytes < 4096 ? 4096 : bytes;
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, sz, hipHostMallocDefault);
    if (e != hipSuccess) {
        *status = hip_check(ctx, e, "pinned allocation");
        return nullptr;
    }
    ctx->pinned = p;
    ctx->pinned_size = sz;
    return p;
}

static hipEvent_t take_event(modp_ctx *ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

TimedLaunch::TimedLaunch(modp_ctx *c, const char *name) : ctx(c), t(nullptr)
{
    if (!ctx->timing || ctx->in_noalloc) return;
    ctx->timings.push_back(KernelTiming{name, take_event(ctx), take_event(ctx)});
    t = &ctx->timings.back();
    (void)hipEventRecord(t->a, ctx->stream);
}

TimedLaunch::~TimedLaunch()
{
    if (t) (void)hipEventRecord(ctx->timings.back().b, ctx->stream);
}

// ---------------------------------------------------------------------------
// argument checking helpers
// ---------------------------------------------------------------------------

static int kind_of(modp_ctx *ctx, int dtype, int *kind)
{
    if (dtype == MODP_F64) { *kind = EK_F64; return MODP_OK; }
    if (dtype == MODP_F32) { *kind = EK_F32; return MODP_OK; }
    return set_error(ctx, MODP_ERR_DTYPE, "unsupported element type %d", dtype);
}

static int check_mat(modp_ctx *ctx, const char *name, const void *X, int64_t rows, int64_t cols, int64_t ld)
{
    if (rows < 0 || cols < 0) return set_error(ctx, MODP_ERR_DIM, "%s: negative dimension", name);
    if (rows > 0 && cols > 0 && !X) return set_error(ctx, MODP_ERR_ARG, "%s: null pointer", name);
    if (ld < (rows > 1 ? rows : 1)) return set_error(ctx, MODP_ERR_ARG, "%s: leading dimension %lld < rows %lld", name, (long long)ld, (long long)rows);
    return MODP_OK;
}

// largest stripe that keeps partial sums exact in the storage type
// (find_max_stripe_ops, stripe_mul.jl:13-27): floor((2^bits-1)/(N-1)^2) - 1
static int check_modulus_gemm(modp_ctx *ctx, int dtype, int64_t p)
{
    if (p < 2) return set_error(ctx, MODP_ERR_MODULUS_RANGE, "modulus %lld < 2", (long long)p);
    const int bits = dtype == MODP_F64 ? 53 : 24;
    const long double top = ldexpl(1.0L, bits) - 1.0L;
    const long double d = (long double)(p - 1) * (long double)(p - 1);
    const long double M = floorl(top / d) - 1.0L;
    if (M < 1.0L)
        return set_error(ctx, MODP_ERR_MODULUS_RANGE,
                         "cannot perform a single multiplication for modulus %lld with datatype %s", (long long)p,
                         dtype == MODP_F64 ? "Float64" : "Float32");
    return MODP_OK;
}

static int check_modulus_elim(modp_ctx *ctx, int dtype, int64_t p)
{
    if (p < 2) return set_error(ctx, MODP_ERR_MODULUS_RANGE, "modulus %lld < 2", (long long)p);
    // residues and pivot-row products must be exact: p <= 2^26 (Float64);
    // Float32 storage holds residues exactly up to 2^24
    const int64_t lim = dtype == MODP_F64 ? ((int64_t)1 << 26) : ((int64_t)1 << 24);
    if (p > lim)
        return set_error(ctx, MODP_ERR_MODULUS_RANGE, "modulus %lld too large for elimination with %s storage",
                         (long long)p, dtype == MODP_F64 ? "Float64" : "Float32");
    return MODP_OK;
}

#define CTX_CHECK(ctx) \
    if (!(ctx)) return MODP_ERR_ARG; \
    (ctx)->last_error.clear();

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

const char *modp_version(void) { return MODP_VERSION; }

int modp_device_count(int *count)
{
    if (!count) return MODP_ERR_ARG;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) { *count = 0; return MODP_ERR_NO_DEVICE; }
    *count = n;
    return MODP_OK;
}

// ---------------------------------------------------------------------------
// the allocation-free product's scratch
// ---------------------------------------------------------------------------

// Digit planes of modp_gemm_noalloc.  Each device holds its own copy, at its
// own address, for the life of the process.  Host code never passes `scratch`
// itself to a kernel (on the host that name is only a stand-in): it uses the
// address hipGetSymbolAddress returns for the device, taken once per device
// when the first context on it is created.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) int8_t scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
int8_t *g_scratch_ptr[MAX_DEVICES] = {nullptr};
hipError_t g_scratch_err[MAX_DEVICES] = {hipSuccess};
bool g_scratch_done[MAX_DEVICES] = {false};

// with `device` current: the device address of its scratch (once per device)
int8_t *scratch_address(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_done[device]) {
        void *ptr = nullptr;
        g_scratch_err[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        g_scratch_ptr[device] = g_scratch_err[device] == hipSuccess ? (int8_t *)ptr : nullptr;
        g_scratch_done[device] = true;
    }
    return g_scratch_ptr[device];
}
}  // namespace

int modp_ctx_create(int device, modp_ctx **out)
{
    if (!out) return MODP_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MODP_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return MODP_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return MODP_ERR_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return MODP_ERR_DEVICE;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return MODP_ERR_NO_DEVICE;
    modp_ctx *c = new modp_ctx();
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    // a failure here leaves only modp_gemm_noalloc unavailable on this device
    c->noalloc_scratch = scratch_address(device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return MODP_ERR_DEVICE;
    }
    *out = c;
    return MODP_OK;
}

int modp_ctx_destroy(modp_ctx *ctx)
{
    if (!ctx) return MODP_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < WS_COUNT; ++i)
        if (ctx->ws[i]) (void)hipFree(ctx->ws[i]);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->invtab) (void)hipFree(ctx->invtab);
    for (auto &t : ctx->timings) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    for (auto e : ctx->slots)
        if (e) (void)hipEventDestroy(e);
    for (int i = 0; i < modp_ctx::NSIDE; ++i) {
        if (ctx->side[i]) (void)hipStreamDestroy(ctx->side[i]);
        if (ctx->join_ev[i]) (void)hipEventDestroy(ctx->join_ev[i]);
    }
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return MODP_OK;
}

const char *modp_last_error(modp_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int modp_synchronize(modp_ctx *ctx)
{
    CTX_CHECK(ctx);
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

void *modp_stream(modp_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int modp_malloc(modp_ctx *ctx, size_t bytes, void **dptr)
{
    CTX_CHECK(ctx);
    if (!dptr) return set_error(ctx, MODP_ERR_ARG, "null output pointer");
    *dptr = nullptr;
    if (bytes == 0) bytes = 1;
    return hip_check(ctx, hipMalloc(dptr, bytes), "hipMalloc");
}

int modp_free(modp_ctx *ctx, void *dptr)
{
    CTX_CHECK(ctx);
    if (!dptr) return MODP_OK;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return hip_check(ctx, hipFree(dptr), "hipFree");
}

int modp_memcpy_h2d(modp_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    CTX_CHECK(ctx);
    if (bytes == 0) return MODP_OK;
    HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipMemcpy H2D");
}

int modp_memcpy_d2h(modp_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    CTX_CHECK(ctx);
    if (bytes == 0) return MODP_OK;
    HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipMemcpy D2H");
}

int modp_memcpy_d2d(modp_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    CTX_CHECK(ctx);
    if (bytes == 0) return MODP_OK;
    return hip_check(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream), "hipMemcpy D2D");
}

int modp_memset_zero(modp_ctx *ctx, void *dptr, size_t bytes)
{
    CTX_CHECK(ctx);
    if (bytes == 0) return MODP_OK;
    return hip_check(ctx, hipMemsetAsync(dptr, 0, bytes, ctx->stream), "hipMemsetAsync");
}

int modp_mod(modp_ctx *ctx, int dtype, int64_t rows, int64_t cols, void *X, int64_t ldx, int64_t N)
{
    CTX_CHECK(ctx);
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    MODP_TRY(check_mat(ctx, "X", X, rows, cols, ldx));
    if (N < 1 || N > ((int64_t)1 << 52))
        return set_error(ctx, MODP_ERR_MODULUS_RANGE, "Modulus %lld is outside 1..2^52", (long long)N);
    return util_mod(ctx, kind, rows, cols, X, ldx, N);
}

int modp_rand(modp_ctx *ctx, int dtype, int64_t rows, int64_t cols, void *X, int64_t ldx, int64_t p, uint64_t seed,
              uint64_t offset)
{
    CTX_CHECK(ctx);
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    MODP_TRY(check_mat(ctx, "X", X, rows, cols, ldx));
    const int64_t lim = dtype == MODP_F64 ? ((int64_t)1 << 52) : ((int64_t)1 << 24);
    if (p < 1 || p > lim) return set_error(ctx, MODP_ERR_MODULUS_RANGE, "modulus %lld out of range", (long long)p);
    return util_rand(ctx, kind, rows, cols, X, ldx, p, seed, offset);
}

int modp_eye(modp_ctx *ctx, int dtype, int64_t rows, int64_t cols, void *X, int64_t ldx)
{
    CTX_CHECK(ctx);
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    MODP_TRY(check_mat(ctx, "X", X, rows, cols, ldx));
    return util_eye(ctx, kind, rows, cols, X, ldx);
}

namespace {

// argument checks shared by modp_gemm and modp_gemm_ex; fills g
int gemm_prepare(modp_ctx *ctx, int dtype, int64_t m, int64_t n, int64_t k, const void *A, int64_t lda,
                 const void *B, int64_t ldb, void *C, int64_t ldc, int64_t p, GemmArgs &g)
{
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    MODP_TRY(check_mat(ctx, "A", A, m, k, lda));
    MODP_TRY(check_mat(ctx, "B", B, k, n, ldb));
    MODP_TRY(check_mat(ctx, "C", C, m, n, ldc));
    MODP_TRY(check_modulus_gemm(ctx, dtype, p));
    if (p > ((int64_t)1 << 26))
        return set_error(ctx, MODP_ERR_MODULUS_RANGE, "modulus %lld above 2^26", (long long)p);
    g.m = m; g.n = n; g.k = k;
    g.A = A; g.lda = lda; g.ka = kind;
    g.B = B; g.ldb = ldb; g.kb = kind;
    g.C = C; g.ldc = ldc; g.kc = kind;
    g.epi = EPI_STORE;
    g.mp = make_modp(p);
    return MODP_OK;
}

// C = 0 for an empty contraction
int gemm_zero(modp_ctx *ctx, const GemmArgs &g)
{
    const size_t es = g.kc == EK_F64 ? 8 : 4;
    return hip_check(ctx, hipMemset2DAsync(g.C, (size_t)g.ldc * es, 0, (size_t)g.m * es, (size_t)g.n, ctx->stream),
                     "hipMemset2DAsync");
}

}  // namespace

int modp_gemm(modp_ctx *ctx, int form, int dtype, int64_t m, int64_t n, int64_t k, const void *A, int64_t lda,
              const void *B, int64_t ldb, void *C, int64_t ldc, int64_t p)
{
    CTX_CHECK(ctx);
    GemmArgs g;
    MODP_TRY(gemm_prepare(ctx, dtype, m, n, k, A, lda, B, ldb, C, ldc, p, g));
    if (form != MODP_FORM_AUTO && form != MODP_FORM_PLAIN_F64 && form != MODP_FORM_LIMB_I8 &&
        form != MODP_FORM_RNS_I8)
        return set_error(ctx, MODP_ERR_ARG, "unknown form %d", form);
    if (m == 0 || n == 0) return MODP_OK;
    if (k == 0) return gemm_zero(ctx, g);
    if (form == MODP_FORM_AUTO) return gemm_auto(ctx, g);
    if (form == MODP_FORM_LIMB_I8) return gemm_limb_i8(ctx, g);
    if (form == MODP_FORM_RNS_I8) return gemm_rns_i8(ctx, g);
    return gemm_plain_f64(ctx, g);
}

int modp_gemm_noalloc(modp_ctx *ctx, int dtype, int64_t m, int64_t n, int64_t k, const void *A, int64_t lda,
                      const void *B, int64_t ldb, void *C, int64_t ldc, int64_t p)
{
    CTX_CHECK(ctx);
    if (!ctx->noalloc_scratch)
        return set_error(ctx, MODP_ERR_DEVICE, "the no-allocation scratch has no device address on device %d",
                         ctx->device);
    GemmArgs g;
    MODP_TRY(gemm_prepare(ctx, dtype, m, n, k, A, lda, B, ldb, C, ldc, p, g));
    if (m == 0 || n == 0) return MODP_OK;
    if (k == 0) return gemm_zero(ctx, g);
    const size_t need = gemm_limb_scratch_bytes(ctx, g);
    if (need > NOALLOC_SCRATCH_BYTES)
        return set_error(ctx, MODP_ERR_ARG,
                         "%lld x %lld x %lld needs %zu bytes of digit planes, above the %zu-byte no-allocation scratch",
                         (long long)m, (long long)n, (long long)k, need, NOALLOC_SCRATCH_BYTES);
    ctx->in_noalloc = true;
    const int st = gemm_limb_i8_in(ctx, g, LimbScratch{ctx->noalloc_scratch, NOALLOC_SCRATCH_BYTES});
    ctx->in_noalloc = false;
    return st;
}

int modp_gemm_ex(modp_ctx *ctx, const modp_gemm_config *cfg, int dtype, int64_t m, int64_t n, int64_t k,
                 const void *A, int64_t lda, const void *B, int64_t ldb, void *C, int64_t ldc, int64_t p)
{
    CTX_CHECK(ctx);
    if (!cfg) return set_error(ctx, MODP_ERR_ARG, "null configuration");
    GemmArgs g;
    MODP_TRY(gemm_prepare(ctx, dtype, m, n, k, A, lda, B, ldb, C, ldc, p, g));
    GemmConfig c;
    c.form = cfg->form; c.tile_m = cfg->tile_m; c.tile_n = cfg->tile_n; c.ksplit = cfg->ksplit;
    if (!gemm_config_valid(g, c))
        return set_error(ctx, MODP_ERR_ARG, "configuration (form %d, tile %dx%d, ksplit %d) invalid for modulus %lld",
                         cfg->form, cfg->tile_m, cfg->tile_n, cfg->ksplit, (long long)p);
    if (m == 0 || n == 0) return MODP_OK;
    if (k == 0) return gemm_zero(ctx, g);
    return gemm_run(ctx, g, c);
}

int modp_autotune_mode(modp_ctx *ctx, int mode)
{
    CTX_CHECK(ctx);
    if (mode < MODP_TUNE_OFF || mode > MODP_TUNE_ON_MISS) return set_error(ctx, MODP_ERR_ARG, "unknown tune mode %d", mode);
    ctx->tune_mode = mode;
    return MODP_OK;
}

int modp_gemm_config_get(modp_ctx *ctx, int dtype, int64_t m, int64_t n, int64_t k, int64_t p,
                         modp_gemm_config *cfg, int *tuned)
{
    CTX_CHECK(ctx);
    if (!cfg) return set_error(ctx, MODP_ERR_ARG, "null output");
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    if (m < 0 || n < 0 || k < 0) return set_error(ctx, MODP_ERR_DIM, "negative size");
    MODP_TRY(check_modulus_gemm(ctx, dtype, p));
    GemmArgs g;
    g.m = m; g.n = n; g.k = k;
    g.ka = g.kb = g.kc = kind;
    g.mp = make_modp(p);
    int t = 0;
    GemmConfig c = gemm_lookup_config(ctx, g, &t);
    cfg->form = c.form; cfg->tile_m = c.tile_m; cfg->tile_n = c.tile_n; cfg->ksplit = c.ksplit;
    if (tuned) *tuned = t;
    return MODP_OK;
}

int modp_autotune_gemm(modp_ctx *ctx, int dtype, int64_t m, int64_t n, int64_t k, int64_t p,
                       modp_gemm_config *best, double *best_ms)
{
    CTX_CHECK(ctx);
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    if (m <= 0 || n <= 0 || k <= 0) return set_error(ctx, MODP_ERR_DIM, "tuning needs positive sizes");
    MODP_TRY(check_modulus_gemm(ctx, dtype, p));
    if (p > ((int64_t)1 << 26)) return set_error(ctx, MODP_ERR_MODULUS_RANGE, "modulus %lld above 2^26", (long long)p);
    const size_t es = kind == EK_F64 ? 8 : 4;
    int st = MODP_OK;
    void *A = ws_get(ctx, WS_TUNE_A, es * (size_t)m * (size_t)k, &st);
    if (!A) return st;
    void *B = ws_get(ctx, WS_TUNE_B, es * (size_t)k * (size_t)n, &st);
    if (!B) return st;
    void *C = ws_get(ctx, WS_TUNE_C, es * (size_t)m * (size_t)n, &st);
    if (!C) return st;
    MODP_TRY(util_rand(ctx, kind, m, k, A, m, p, 0x7475ULL, 0));
    MODP_TRY(util_rand(ctx, kind, k, n, B, k, p, 0x7475ULL, (uint64_t)(m * k)));
    GemmArgs g;
    g.m = m; g.n = n; g.k = k;
    g.A = A; g.lda = m; g.ka = kind;
    g.B = B; g.ldb = k; g.kb = kind;
    g.C = C; g.ldc = m; g.kc = kind;
    g.epi = EPI_STORE;
    g.mp = make_modp(p);
    GemmConfig c;
    double ms = 0;
    MODP_TRY(tune_gemm(ctx, g, &c, &ms));
    if (best) { best->form = c.form; best->tile_m = c.tile_m; best->tile_n = c.tile_n; best->ksplit = c.ksplit; }
    if (best_ms) *best_ms = ms;
    return MODP_OK;
}

int modp_autotune_export(modp_ctx *ctx, char *buf, size_t cap, size_t *len)
{
    CTX_CHECK(ctx);
    return tune_export(ctx, buf, cap, len);
}

int modp_autotune_import(modp_ctx *ctx, const char *text)
{
    CTX_CHECK(ctx);
    return tune_import(ctx, text);
}

int modp_pluq(modp_ctx *ctx, int form, int dtype, int64_t m, int64_t n, const void *A, int64_t lda, void *U,
              int64_t ldu, void *L, int64_t ldl, int64_t p, int64_t *row_swaps, int64_t *n_row, int64_t *col_swaps,
              int64_t *n_col, int64_t *rank)
{
    CTX_CHECK(ctx);
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    MODP_TRY(check_mat(ctx, "A", A, m, n, lda));
    MODP_TRY(check_mat(ctx, "U", U, m, n, ldu));
    MODP_TRY(check_mat(ctx, "L", L, m, m, ldl));
    MODP_TRY(check_modulus_elim(ctx, dtype, p));
    if (!row_swaps || !n_row || !col_swaps || !n_col || !rank) return set_error(ctx, MODP_ERR_ARG, "null output");
    ElimResult r;
    MODP_TRY(elim_run(ctx, ELIM_PLUQ, form, kind, m, n, A, lda, U, ldu, L, ldl, p, r));
    std::copy(r.row_swaps.begin(), r.row_swaps.end(), row_swaps);
    std::copy(r.col_swaps.begin(), r.col_swaps.end(), col_swaps);
    *n_row = (int64_t)r.row_swaps.size() / 2;
    *n_col = (int64_t)r.col_swaps.size() / 2;
    *rank = r.rank;
    return MODP_OK;
}

int modp_lu(modp_ctx *ctx, int form, int dtype, int64_t m, int64_t n, const void *A, int64_t lda, void *U, int64_t ldu,
            void *L, int64_t ldl, int64_t p, int64_t *perm, int64_t *pivcols, int64_t *rank)
{
    CTX_CHECK(ctx);
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    MODP_TRY(check_mat(ctx, "A", A, m, n, lda));
    MODP_TRY(check_mat(ctx, "U", U, m, n, ldu));
    MODP_TRY(check_mat(ctx, "L", L, m, m, ldl));
    MODP_TRY(check_modulus_elim(ctx, dtype, p));
    if ((!perm && m > 0) || !rank) return set_error(ctx, MODP_ERR_ARG, "null output");
    ElimResult r;
    MODP_TRY(elim_run(ctx, ELIM_LU, form, kind, m, n, A, lda, U, ldu, L, ldl, p, r));
    std::copy(r.perm.begin(), r.perm.end(), perm);
    if (pivcols) std::copy(r.pivcols.begin(), r.pivcols.end(), pivcols);
    *rank = r.rank;
    return MODP_OK;
}

int modp_rref(modp_ctx *ctx, int form, int dtype, int64_t m, int64_t n, const void *A, int64_t lda, void *R,
              int64_t ldr, int64_t p, int64_t *pivcols, int64_t *rank)
{
    CTX_CHECK(ctx);
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    MODP_TRY(check_mat(ctx, "A", A, m, n, lda));
    MODP_TRY(check_mat(ctx, "R", R, m, n, ldr));
    MODP_TRY(check_modulus_elim(ctx, dtype, p));
    if (!rank) return set_error(ctx, MODP_ERR_ARG, "null output");
    ElimResult r;
    if (m == 0 || n == 0) { *rank = 0; return MODP_OK; }
    MODP_TRY(elim_rref(ctx, form, kind, m, n, A, lda, R, ldr, p, r));
    if (pivcols) std::copy(r.pivcols.begin(), r.pivcols.end(), pivcols);
    *rank = r.rank;
    return MODP_OK;
}

int modp_apply_row_perm(modp_ctx *ctx, int dtype, int64_t rows, int64_t cols, void *X, int64_t ldx,
                        const int64_t *stack, int64_t nswaps, int inverse)
{
    CTX_CHECK(ctx);
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    MODP_TRY(check_mat(ctx, "X", X, rows, cols, ldx));
    if (nswaps > 0 && !stack) return set_error(ctx, MODP_ERR_ARG, "null stack");
    return util_perm(ctx, kind, 1, rows, cols, X, ldx, stack, nswaps, inverse);
}

int modp_apply_col_perm(modp_ctx *ctx, int dtype, int64_t rows, int64_t cols, void *X, int64_t ldx,
                        const int64_t *stack, int64_t nswaps, int inverse)
{
    CTX_CHECK(ctx);
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    MODP_TRY(check_mat(ctx, "X", X, rows, cols, ldx));
    if (nswaps > 0 && !stack) return set_error(ctx, MODP_ERR_ARG, "null stack");
    return util_perm(ctx, kind, 0, rows, cols, X, ldx, stack, nswaps, inverse);
}

// ---- batched ----

int modp_gemm_batched_plan(modp_ctx *ctx, int dtype, const modp_gemm_job *jobs, int64_t count, modp_gemm_plan **out)
{
    CTX_CHECK(ctx);
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    if (!out || (count > 0 && !jobs)) return set_error(ctx, MODP_ERR_ARG, "null argument");
    for (int64_t i = 0; i < count; ++i) {
        const modp_gemm_job &j = jobs[i];
        MODP_TRY(check_mat(ctx, "A", j.A, j.m, j.k, j.lda));
        MODP_TRY(check_mat(ctx, "B", j.B, j.k, j.n, j.ldb));
        MODP_TRY(check_mat(ctx, "C", j.C, j.m, j.n, j.ldc));
        MODP_TRY(check_modulus_gemm(ctx, dtype, j.p));
        if (j.p > ((int64_t)1 << 26))
            return set_error(ctx, MODP_ERR_MODULUS_RANGE, "job %lld: modulus above 2^26", (long long)i);
    }
    return batched_gemm_plan(ctx, kind, jobs, count, out);
}

int modp_gemm_batched_run(modp_ctx *ctx, modp_gemm_plan *plan)
{
    CTX_CHECK(ctx);
    if (!plan) return set_error(ctx, MODP_ERR_ARG, "null plan");
    return batched_gemm_run(ctx, plan);
}

int modp_gemm_batched_destroy(modp_ctx *ctx, modp_gemm_plan *plan)
{
    CTX_CHECK(ctx);
    return batched_gemm_destroy(ctx, plan);
}

int modp_gemm_batched(modp_ctx *ctx, int dtype, const modp_gemm_job *jobs, int64_t count)
{
    modp_gemm_plan *plan = nullptr;
    MODP_TRY(modp_gemm_batched_plan(ctx, dtype, jobs, count, &plan));
    int st = modp_gemm_batched_run(ctx, plan);
    int st2 = modp_gemm_batched_destroy(ctx, plan);
    return st != MODP_OK ? st : st2;
}

static int check_elim_jobs(modp_ctx *ctx, int op, int dtype, const modp_elim_job *jobs, int64_t count, int *kind)
{
    MODP_TRY(kind_of(ctx, dtype, kind));
    if (op != MODP_ELIM_PLUQ && op != MODP_ELIM_LU && op != MODP_ELIM_RREF)
        return set_error(ctx, MODP_ERR_ARG, "unknown elimination op %d", op);
    if (count > 0 && !jobs) return set_error(ctx, MODP_ERR_ARG, "null jobs");
    for (int64_t i = 0; i < count; ++i) {
        const modp_elim_job &j = jobs[i];
        if (j.m < 0 || j.n < 0 || j.m > 128 || j.n > 128)
            return set_error(ctx, MODP_ERR_DIM, "job %lld: batched elimination needs m, n <= 128", (long long)i);
        MODP_TRY(check_mat(ctx, "A", j.A, j.m, j.n, j.lda));
        MODP_TRY(check_mat(ctx, "U", j.U, j.m, j.n, j.ldu));
        if (op != MODP_ELIM_RREF) MODP_TRY(check_mat(ctx, "L", j.L, j.m, j.m, j.ldl));
        MODP_TRY(check_modulus_elim(ctx, dtype, j.p));
        if (!j.counts || !j.row_info || (op != MODP_ELIM_RREF && !j.col_info))
            return set_error(ctx, MODP_ERR_ARG, "job %lld: null output", (long long)i);
    }
    return MODP_OK;
}

int modp_elim_batched(modp_ctx *ctx, int op, int dtype, const modp_elim_job *jobs, int64_t count)
{
    CTX_CHECK(ctx);
    int kind;
    MODP_TRY(check_elim_jobs(ctx, op, dtype, jobs, count, &kind));
    return batched_elim(ctx, op, kind, jobs, count);
}

int modp_elim_batched_plan(modp_ctx *ctx, int op, int dtype, const modp_elim_job *jobs, int64_t count,
                           modp_elim_plan **out)
{
    CTX_CHECK(ctx);
    if (!out) return set_error(ctx, MODP_ERR_ARG, "null output");
    *out = nullptr;
    int kind;
    MODP_TRY(check_elim_jobs(ctx, op, dtype, jobs, count, &kind));
    return batched_elim_plan(ctx, op, kind, jobs, count, out);
}

int modp_elim_batched_run(modp_ctx *ctx, modp_elim_plan *plan)
{
    CTX_CHECK(ctx);
    if (!plan) return set_error(ctx, MODP_ERR_ARG, "null plan");
    return batched_elim_run(ctx, plan);
}

int modp_elim_batched_destroy(modp_ctx *ctx, modp_elim_plan *plan)
{
    CTX_CHECK(ctx);
    return batched_elim_destroy(ctx, plan);
}

// ---- timing ----

int modp_kernel_timing(modp_ctx *ctx, int enable)
{
    CTX_CHECK(ctx);
    ctx->timing = enable != 0;
    return MODP_OK;
}

int modp_kernel_timing_reset(modp_ctx *ctx)
{
    CTX_CHECK(ctx);
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (auto &t : ctx->timings) {
        ctx->event_pool.push_back(t.a);
        ctx->event_pool.push_back(t.b);
    }
    ctx->timings.clear();
    return MODP_OK;
}

int modp_kernel_time(modp_ctx *ctx, const char *prefix, double *total_ms, int64_t *count)
{
    CTX_CHECK(ctx);
    if (!prefix || !total_ms || !count) return set_error(ctx, MODP_ERR_ARG, "null argument");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    double tot = 0.0;
    int64_t cnt = 0;
    const size_t pl = strlen(prefix);
    for (auto &t : ctx->timings) {
        if (t.name.compare(0, pl, prefix) != 0) continue;
        float ms = 0.f;
        HIP_TRY(ctx, hipEventElapsedTime(&ms, t.a, t.b));
        tot += ms;
        ++cnt;
    }
    *total_ms = tot;
    *count = cnt;
    return MODP_OK;
}

int modp_event_record(modp_ctx *ctx, int slot)
{
    CTX_CHECK(ctx);
    if (slot < 0 || slot >= 16) return set_error(ctx, MODP_ERR_ARG, "slot out of range");
    if (!ctx->slots[slot]) HIP_TRY(ctx, hipEventCreate(&ctx->slots[slot]));
    return hip_check(ctx, hipEventRecord(ctx->slots[slot], ctx->stream), "hipEventRecord");
}

int modp_event_elapsed(modp_ctx *ctx, int a, int b, float *ms)
{
    CTX_CHECK(ctx);
    if (a < 0 || a >= 16 || b < 0 || b >= 16 || !ctx->slots[a] || !ctx->slots[b] || !ms)
        return set_error(ctx, MODP_ERR_ARG, "bad event slots");
    HIP_TRY(ctx, hipEventSynchronize(ctx->slots[b]));
    return hip_check(ctx, hipEventElapsedTime(ms, ctx->slots[a], ctx->slots[b]), "hipEventElapsedTime");
}

}  // extern "C"

int modp_probe_peak(modp_ctx *ctx, int kind, double *rate)
{
    CTX_CHECK(ctx);
    if (!rate) return set_error(ctx, MODP_ERR_ARG, "null output");
    if (kind < 0 || kind > 3) return set_error(ctx, MODP_ERR_ARG, "unknown peak kind %d", kind);
    return probe_peak(ctx, kind, rate);
}

int modp_trtri(modp_ctx *ctx, int dtype, int uplo, int64_t m, int64_t n, const void *A, int64_t lda, void *Ainv,
               int64_t ldai, int64_t p)
{
    CTX_CHECK(ctx);
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    if (uplo != 0 && uplo != 1) return set_error(ctx, MODP_ERR_ARG, "uplo must be 0 (upper) or 1 (lower)");
    MODP_TRY(check_mat(ctx, "A", A, m, n, lda));
    MODP_TRY(check_mat(ctx, "Ainv", Ainv, n, m, ldai));
    MODP_TRY(check_modulus_elim(ctx, dtype, p));
    return tri_inverse(ctx, kind, uplo == 0, m, n, A, lda, Ainv, ldai, p);
}

int modp_inverse(modp_ctx *ctx, int form, int dtype, int64_t n, const void *A, int64_t lda, void *Ainv, int64_t ldai,
                 int64_t p, int *invertible)
{
    CTX_CHECK(ctx);
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    if (!invertible) return set_error(ctx, MODP_ERR_ARG, "null output");
    if (form != MODP_FORM_AUTO && form != MODP_FORM_PLAIN_F64 && form != MODP_FORM_LIMB_I8)
        return set_error(ctx, MODP_ERR_ARG, "unknown form %d", form);
    MODP_TRY(check_mat(ctx, "A", A, n, n, lda));
    MODP_TRY(check_mat(ctx, "Ainv", Ainv, n, n, ldai));
    MODP_TRY(check_modulus_elim(ctx, dtype, p));
    return mat_inverse(ctx, form, kind, n, A, lda, Ainv, ldai, p, invertible);
}

int modp_gemv(modp_ctx *ctx, int dtype, int64_t m, int64_t k, const void *A, int64_t lda, const void *x, void *y,
              int64_t p)
{
    CTX_CHECK(ctx);
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    MODP_TRY(check_mat(ctx, "A", A, m, k, lda));
    if ((k > 0 && !x) || (m > 0 && !y)) return set_error(ctx, MODP_ERR_ARG, "null vector");
    MODP_TRY(check_modulus_gemm(ctx, dtype, p));
    if (p > ((int64_t)1 << 26))
        return set_error(ctx, MODP_ERR_MODULUS_RANGE, "modulus %lld above 2^26", (long long)p);
    return gemv_run(ctx, kind, m, k, A, lda, x, y, p);
}

int modp_crt(modp_ctx *ctx, int dtype, int64_t rows, int64_t cols, int r, const void *const *X,
             const int64_t *ldx, const int64_t *p, int sign, uint32_t *Y, int64_t ldy, int64_t words)
{
    CTX_CHECK(ctx);
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    if (r < 1 || r > 12 || !X || !ldx || !p) return set_error(ctx, MODP_ERR_ARG, "CRT needs 1..12 residue matrices");
    for (int i = 0; i < r; ++i) MODP_TRY(check_mat(ctx, "X", X[i], rows, cols, ldx[i]));
    if (rows > 0 && cols > 0 && (!Y || ldy < rows))
        return set_error(ctx, MODP_ERR_ARG, "CRT output: null pointer or leading dimension below the rows");
    return util_crt(ctx, kind, rows, cols, r, X, ldx, p, sign, Y, ldy, words);
}

int modp_crt2(modp_ctx *ctx, int dtype, int64_t rows, int64_t cols, const void *X1, int64_t ld1, int64_t p1,
              const void *X2, int64_t ld2, int64_t p2, void *Y, int64_t ldy)
{
    CTX_CHECK(ctx);
    int kind;
    MODP_TRY(kind_of(ctx, dtype, &kind));
    MODP_TRY(check_mat(ctx, "X1", X1, rows, cols, ld1));
    MODP_TRY(check_mat(ctx, "X2", X2, rows, cols, ld2));
    MODP_TRY(check_mat(ctx, "Y", Y, rows, cols, ldy));
    const int64_t lim26 = (int64_t)1 << 26;
    if (p1 < 2 || p2 < 2 || p1 > lim26 || p2 > lim26)
        return set_error(ctx, MODP_ERR_MODULUS_RANGE, "moduli %lld, %lld outside 2..2^26", (long long)p1, (long long)p2);
    const int64_t lim = dtype == MODP_F64 ? ((int64_t)1 << 53) : ((int64_t)1 << 24);
    if (p1 * p2 > lim)
        return set_error(ctx, MODP_ERR_MODULUS_RANGE, "p1*p2 = %lld is not exact in the storage type",
                         (long long)(p1 * p2));
    return util_crt2(ctx, kind, rows, cols, X1, ld1, p1, X2, ld2, p2, Y, ldy);
}
