// peak.hip -- microbenchmarks of the device peaks that bound the hot path:
// dense int8 MFMA (v_mfma_i32_16x16x64_i8 and v_mfma_i32_32x32x32_i8), FP64
// MFMA (v_mfma_f64_16x16x4f64) and HBM streaming bandwidth.  Used by the benchmark to report roofline
// fractions against measured as well as published peaks.
#include "common.h"

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

constexpr int CHAINS = 8;  // independent accumulators per wave: covers MFMA latency

__global__ __launch_bounds__(256) void k_peak_i8(int iters, int seed, int *out)
{
    const int lane = threadIdx.x & 63;
    v4i a = v4i{seed + lane, lane, 3, seed};
    v4i b = v4i{lane ^ seed, 7, seed, lane};
    v4i acc[CHAINS];
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) acc[c] = v4i{c, 0, 0, 0};
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < CHAINS; ++c) acc[c] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, acc[c], 0, 0, 0);
    }
    int s = 0;
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_peak_i8_32(int iters, int seed, int *out)
{
    const int lane = threadIdx.x & 63;
    v4i a = v4i{seed + lane, lane, 3, seed};
    v4i b = v4i{lane ^ seed, 7, seed, lane};
    v16i acc[CHAINS];
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) acc[c] = v16i{c};
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < CHAINS; ++c) acc[c] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc[c], 0, 0, 0);
    }
    int s = 0;
#pragma unroll
    for (int c = 0; c < CHAINS; ++c)
#pragma unroll
        for (int e = 0; e < 16; ++e) s += acc[c][e];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_peak_f64(int iters, double seed, double *out)
{
    const int lane = threadIdx.x & 63;
    const double a = seed + lane, b = 1.0 - seed * lane;
    v4d acc[CHAINS];
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) acc[c] = v4d{(double)c, 0.0, 0.0, 0.0};
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < CHAINS; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
    }
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// y = x + 1 over n 16-byte words (read + write stream)
__global__ __launch_bounds__(256) void k_peak_copy(const v4i *__restrict__ x, v4i *__restrict__ y, int64_t n)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        v4i v = x[i];
        y[i] = v + 1;
    }
}

}  // namespace

int probe_peak(modp_ctx *ctx, int kind, double *rate)
{
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(ctx, hipEventCreate(&a));
    int st = hip_check(ctx, hipEventCreate(&b), "hipEventCreate");
    const int grid = ctx->num_cus * 8;  // 8 workgroups of 4 waves per CU
    void *buf = nullptr;
    double work = 0.0;
    float best = 1e30f;
    if (st == MODP_OK) {
        size_t bytes = 0;
        if (kind == 2) bytes = (size_t)2 << 30;  // 1 GiB in, 1 GiB out
        else bytes = (size_t)grid * 256 * 8;
        buf = ws_get(ctx, WS_SCRATCH, bytes, &st);
    }
    const int iters = 4096;
    for (int r = 0; r < 4 && st == MODP_OK; ++r) {
        st = hip_check(ctx, hipEventRecord(a, ctx->stream), "hipEventRecord");
        if (st != MODP_OK) break;
        if (kind == 0) {
            k_peak_i8<<<grid, 256, 0, ctx->stream>>>(iters, r, (int *)buf);
            work = (double)grid * 4 * iters * CHAINS * (16.0 * 16 * 64 * 2);
        } else if (kind == 3) {
            k_peak_i8_32<<<grid, 256, 0, ctx->stream>>>(iters / 2, r, (int *)buf);
            work = (double)grid * 4 * (iters / 2) * CHAINS * (32.0 * 32 * 32 * 2);
        } else if (kind == 1) {
            k_peak_f64<<<grid, 256, 0, ctx->stream>>>(iters / 4, 0.5 + r, (double *)buf);
            work = (double)grid * 4 * (iters / 4) * CHAINS * (16.0 * 16 * 4 * 2);
        } else {
            const int64_t words = ((int64_t)1 << 30) / 16;
            k_peak_copy<<<ctx->num_cus * 16, 256, 0, ctx->stream>>>((const v4i *)buf,
                                                                  (v4i *)((char *)buf + ((size_t)1 << 30)), words);
            work = 2.0 * (double)((int64_t)1 << 30);
        }
        st = hip_check(ctx, hipGetLastError(), "peak probe launch");
        if (st == MODP_OK) st = hip_check(ctx, hipEventRecord(b, ctx->stream), "hipEventRecord");
        if (st == MODP_OK) st = hip_check(ctx, hipEventSynchronize(b), "hipEventSynchronize");
        float t = 0.f;
        if (st == MODP_OK) st = hip_check(ctx, hipEventElapsedTime(&t, a, b), "hipEventElapsedTime");
        if (st == MODP_OK && r > 0 && t < best) best = t;  // first run warms up
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (st == MODP_OK) *rate = work / (best * 1e-3);  // ops/s or bytes/s
    return st;
}
