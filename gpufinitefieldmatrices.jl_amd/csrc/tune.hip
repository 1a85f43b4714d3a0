// tune.hip -- autotuner for the GEMM launch configuration.
//
// A configuration is (form, output tile, K slices).  The candidates valid for
// a product are timed on the device with HIP events and the fastest is kept
// in the context's cache under a key of (element kinds, bits of p, log2
// buckets of m, n, k); later AUTO-form products of the same bucket reuse it.
// Batched plans are tuned over their 64-tile threshold in batched.hip and
// share the cache (keys with kind + 16).  The cache can be exported as text
// and imported into another context.
#include <algorithm>
#include <cstdio>
#include <cstring>

#include "common.h"

namespace {

int ceil_log2(int64_t x)
{
    int b = 0;
    while (((int64_t)1 << b) < x && b < 62) ++b;
    return b;
}

int bits_of(int64_t p)
{
    int b = 0;
    int64_t x = p - 1;
    while (x > 0) { ++b; x >>= 1; }
    return b;
}

// launch-time of one configuration: one untimed run, then the minimum over
// `reps` timed runs (a single run once it exceeds 50 ms)
int time_config(modp_ctx *ctx, const GemmArgs &g, const GemmConfig &cfg, double *ms)
{
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(ctx, hipEventCreate(&a));
    int st = hip_check(ctx, hipEventCreate(&b), "hipEventCreate");
    double best = 1e30;
    if (st == MODP_OK) st = gemm_run(ctx, g, cfg);
    int reps = 3;
    for (int r = 0; r < reps && st == MODP_OK; ++r) {
        st = hip_check(ctx, hipEventRecord(a, ctx->stream), "hipEventRecord");
        if (st == MODP_OK) st = gemm_run(ctx, g, cfg);
        if (st == MODP_OK) st = hip_check(ctx, hipEventRecord(b, ctx->stream), "hipEventRecord");
        if (st == MODP_OK) st = hip_check(ctx, hipEventSynchronize(b), "hipEventSynchronize");
        float t = 0.f;
        if (st == MODP_OK) st = hip_check(ctx, hipEventElapsedTime(&t, a, b), "hipEventElapsedTime");
        if (st == MODP_OK) {
            best = std::min(best, (double)t);
            if (t > 50.f) reps = 1;
        }
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    *ms = best;
    return st;
}

std::vector<GemmConfig> candidates(const modp_ctx *ctx, const GemmArgs &g)
{
    std::vector<GemmConfig> out;
    const int forms[3] = {MODP_FORM_PLAIN_F64, MODP_FORM_LIMB_I8, MODP_FORM_RNS_I8};
    const int tiles[4][2] = {{64, 64}, {128, 64}, {128, 128}, {256, 256}};
    const int splits[4] = {1, 2, 4, 8};
    for (int f : forms)
        for (const auto &t : tiles)
            for (int s : splits) {
                GemmConfig c;
                c.form = f; c.tile_m = t[0]; c.tile_n = t[1]; c.ksplit = s;
                if (!gemm_config_valid(g, c)) continue;
                const int64_t nt = ((g.m + t[0] - 1) / t[0]) * ((g.n + t[1] - 1) / t[1]);
                // K slices only pay while the grid leaves CUs idle
                if (s > 1 && (nt >= 2 * ctx->num_cus || g.k / s < 256)) continue;
                out.push_back(c);
            }
    return out;
}

}  // namespace

TuneKey tune_key_gemm(const GemmArgs &g)
{
    TuneKey k;
    k.v[0] = g.ka;
    k.v[1] = g.kb;
    k.v[2] = g.kc;
    k.v[3] = bits_of(g.mp.p);
    k.v[4] = ceil_log2(g.m);
    k.v[5] = ceil_log2(g.n);
    k.v[6] = ceil_log2(g.k);
    return k;
}

GemmConfig gemm_lookup_config(const modp_ctx *ctx, const GemmArgs &g, int *tuned)
{
    if (tuned) *tuned = 0;
    if (ctx->tune_mode != MODP_TUNE_OFF) {
        auto it = ctx->tune_cache.find(tune_key_gemm(g));
        if (it != ctx->tune_cache.end() && gemm_config_valid(g, it->second.cfg)) {
            if (tuned) *tuned = 1;
            return it->second.cfg;
        }
    }
    return gemm_default_config(ctx, g, MODP_FORM_AUTO);
}

int tune_gemm(modp_ctx *ctx, const GemmArgs &g, GemmConfig *best, double *best_ms)
{
    std::vector<GemmConfig> cands = candidates(ctx, g);
    if (cands.empty()) return set_error(ctx, MODP_ERR_MODULUS_RANGE, "no GEMM form holds modulus %lld", (long long)g.mp.p);
    double bt = 1e30;
    GemmConfig bc = cands[0];
    for (const GemmConfig &c : cands) {
        double t = 0;
        MODP_TRY(time_config(ctx, g, c, &t));
        if (t < bt) { bt = t; bc = c; }
    }
    TuneEntry e;
    e.cfg = bc;
    e.ms = bt;
    ctx->tune_cache[tune_key_gemm(g)] = e;
    *best = bc;
    if (best_ms) *best_ms = bt;
    return MODP_OK;
}

int gemm_auto(modp_ctx *ctx, const GemmArgs &g)
{
    int tuned = 0;
    GemmConfig cfg = gemm_lookup_config(ctx, g, &tuned);
    // tuning reruns the product, so only a plain store may be tuned in place
    if (!tuned && ctx->tune_mode == MODP_TUNE_ON_MISS && g.epi == EPI_STORE && !g.skip && g.k > 0) {
        double ms = 0;
        MODP_TRY(tune_gemm(ctx, g, &cfg, &ms));
    }
    return gemm_run(ctx, g, cfg);
}

// ---- text form of the cache: one entry per line -------------------------
//   g <ka> <kb> <kc> <bits> <lm> <ln> <lk> <form> <tile_m> <tile_n> <ksplit> <ms>
//   b <kind> <lcount> <lm> <ln> <lk> 0 0 <thr> <ms>

int tune_export(modp_ctx *ctx, char *buf, size_t cap, size_t *len)
{
    std::string out;
    char line[256];
    for (const auto &kv : ctx->tune_cache) {
        const TuneKey &k = kv.first;
        const TuneEntry &e = kv.second;
        if (k.v[0] >= 16)
            snprintf(line, sizeof line, "b %d %d %d %d %d %d %d %.6g %.6g\n", k.v[0] - 16, k.v[1], k.v[2], k.v[3],
                     k.v[4], k.v[5], k.v[6], e.thr, e.ms);
        else
            snprintf(line, sizeof line, "g %d %d %d %d %d %d %d %d %d %d %d %.6g\n", k.v[0], k.v[1], k.v[2],
                     k.v[3], k.v[4], k.v[5], k.v[6], e.cfg.form, e.cfg.tile_m, e.cfg.tile_n, e.cfg.ksplit, e.ms);
        out += line;
    }
    if (len) *len = out.size() + 1;
    if (!buf) return MODP_OK;
    if (cap < out.size() + 1) return set_error(ctx, MODP_ERR_ARG, "buffer of %zu bytes holds less than %zu", cap, out.size() + 1);
    memcpy(buf, out.c_str(), out.size() + 1);
    return MODP_OK;
}

int tune_import(modp_ctx *ctx, const char *text)
{
    if (!text) return set_error(ctx, MODP_ERR_ARG, "null text");
    const char *p = text;
    int lineno = 0;
    while (*p) {
        const char *eol = strchr(p, '\n');
        std::string line(p, eol ? (size_t)(eol - p) : strlen(p));
        p = eol ? eol + 1 : p + line.size();
        ++lineno;
        if (line.empty() || line[0] == '#') continue;
        TuneKey k;
        TuneEntry e;
        if (line[0] == 'g') {
            if (sscanf(line.c_str(), "g %d %d %d %d %d %d %d %d %d %d %d %lf", &k.v[0], &k.v[1], &k.v[2], &k.v[3],
                       &k.v[4], &k.v[5], &k.v[6], &e.cfg.form, &e.cfg.tile_m, &e.cfg.tile_n, &e.cfg.ksplit,
                       &e.ms) != 12 ||
                k.v[0] < 0 || k.v[0] > 2 ||
                (e.cfg.form != MODP_FORM_PLAIN_F64 && e.cfg.form != MODP_FORM_LIMB_I8 && e.cfg.form != MODP_FORM_RNS_I8) ||
                (e.cfg.tile_m != 64 && e.cfg.tile_m != 128 && e.cfg.tile_m != 256) ||
                (e.cfg.tile_n != 64 && e.cfg.tile_n != 128 && e.cfg.tile_n != 256) ||
                e.cfg.ksplit < 1 || e.cfg.ksplit > 16)
                return set_error(ctx, MODP_ERR_ARG, "autotune cache line %d malformed", lineno);
        } else if (line[0] == 'b') {
            if (sscanf(line.c_str(), "b %d %d %d %d %d %d %d %lf %lf", &k.v[0], &k.v[1], &k.v[2], &k.v[3], &k.v[4],
                       &k.v[5], &k.v[6], &e.thr, &e.ms) != 9 || k.v[0] < 0 || k.v[0] > 2)
                return set_error(ctx, MODP_ERR_ARG, "autotune cache line %d malformed", lineno);
            k.v[0] += 16;
        } else {
            return set_error(ctx, MODP_ERR_ARG, "autotune cache line %d malformed", lineno);
        }
        ctx->tune_cache[k] = e;
    }
    return MODP_OK;
}
