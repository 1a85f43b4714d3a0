// batched.h -- batched multi-modular forms (internal interface).
#pragma once
#include "common.h"

int batched_gemm_plan(modp_ctx *ctx, int kind, const modp_gemm_job *jobs, int64_t count, modp_gemm_plan **out);
int batched_gemm_run(modp_ctx *ctx, modp_gemm_plan *plan);
int batched_gemm_destroy(modp_ctx *ctx, modp_gemm_plan *plan);
int batched_elim(modp_ctx *ctx, int op, int kind, const modp_elim_job *jobs, int64_t count);
int batched_elim_plan(modp_ctx *ctx, int op, int kind, const modp_elim_job *jobs, int64_t count,
                      modp_elim_plan **out);
int batched_elim_run(modp_ctx *ctx, modp_elim_plan *plan);
int batched_elim_destroy(modp_ctx *ctx, modp_elim_plan *plan);
