/*
 * modp.h -- C ABI of libmodp_hip.so, the MI355X (gfx950) implementation of
 * GPUFiniteFieldMatrices.jl's modular matrix multiply and mod-p elimination.
 *
 * Every entry point takes plain pointers and sizes.  Matrices are column-major
 * with an explicit leading dimension, the layout of CuModArray's padded
 * `data` array (src/CuModMatrix/CuModMatrix.jl:42-122: storage is
 * (rows+32) x (cols+32), logical size `size`, modulus `N`).  Matrix pointers
 * are device pointers (obtained from modp_malloc or any HIP allocation on the
 * context's device); index stacks, permutations and counts are host pointers.
 * Calls are ordered on the context's stream; entry points that return host
 * values (stacks, rank) synchronise before returning.
 *
 * Status codes map one-to-one onto the reference's exceptions in the host
 * bindings (see modp_status).
 */
#ifndef MODP_H
#define MODP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct modp_ctx modp_ctx;

typedef enum {
    MODP_OK = 0,
    MODP_ERR_DIM = 1,              /* CuModArraySizeMismatchException / DimensionMismatch */
    MODP_ERR_MODULUS_RANGE = 2,    /* modulus outside what the arithmetic can hold exactly
                                      (ArgumentError "cannot perform a single multiplication",
                                      stripe_mul.jl:200-202; N > 2^52, CuModMatrix.jl:55-59) */
    MODP_ERR_MODULUS_MISMATCH = 3, /* CuModArrayModulusMismatchException (CuModMatrix.jl:769-773) */
    MODP_ERR_NOT_INVERTIBLE = 4,   /* a pivot has no inverse mod N: N is not prime
                                      (CuModMatrixModulusNotPrimeException, CuModMatrix.jl:21-23) */
    MODP_ERR_DTYPE = 5,            /* element type unsupported (stripe_mul.jl:181-182) */
    MODP_ERR_OOM = 6,
    MODP_ERR_DEVICE = 7,           /* HIP runtime error */
    MODP_ERR_ARG = 8,              /* invalid argument (null pointer, ld < rows, ...) */
    MODP_ERR_NO_DEVICE = 9
} modp_status;

/* element type of the padded storage (CuModArray{T}: Float64 / Float32) */
typedef enum { MODP_F64 = 0, MODP_F32 = 1 } modp_dtype;

/* arithmetic form of the GEMM and of the elimination's trailing update */
typedef enum {
    MODP_FORM_AUTO = 0,      /* autotuner's choice for (shape, bits of p) */
    MODP_FORM_PLAIN_F64 = 1, /* FP64 MFMA, balanced residues, periodic reduction */
    MODP_FORM_LIMB_I8 = 2,   /* signed base-256 digits (limbs) in int8,
                                int32-accumulating MFMA, FP64 Barrett recombination */
    MODP_FORM_RNS_I8 = 3     /* multi-modular limbs: residues modulo up to 12 pairwise
                                coprime moduli <= 256 in int8, one int32-accumulating
                                MFMA product per modulus, Chinese-remainder
                                recombination in FP64 (GEMM only); for
                                2 <= p <= 255 one product modulo p itself */
} modp_form;

/* autotuner modes (modp_autotune_mode) */
typedef enum {
    MODP_TUNE_OFF = 0,     /* heuristics only (default) */
    MODP_TUNE_CACHED = 1,  /* use tuned configurations present in the cache */
    MODP_TUNE_ON_MISS = 2  /* also tune, on the call's own operands, a shape bucket
                              missing from the cache */
} modp_tune_mode;

/* ---- version, context, errors ---------------------------------------- */

const char *modp_version(void);
int modp_device_count(int *count);
int modp_ctx_create(int device, modp_ctx **out);
int modp_ctx_destroy(modp_ctx *ctx);
/* message of the last failing call on this context ("" if none) */
const char *modp_last_error(modp_ctx *ctx);
int modp_synchronize(modp_ctx *ctx);
/* the context's hipStream_t, for callers that order their own work on it */
void *modp_stream(modp_ctx *ctx);

/* ---- device memory (CuArray allocation / copyto! / Array) ------------- */

int modp_malloc(modp_ctx *ctx, size_t bytes, void **dptr);
int modp_free(modp_ctx *ctx, void *dptr);
int modp_memcpy_h2d(modp_ctx *ctx, void *dst, const void *src, size_t bytes);
int modp_memcpy_d2h(modp_ctx *ctx, void *dst, const void *src, size_t bytes);
int modp_memcpy_d2d(modp_ctx *ctx, void *dst, const void *src, size_t bytes);
int modp_memset_zero(modp_ctx *ctx, void *dptr, size_t bytes);

/* ---- elementwise ------------------------------------------------------ */

/* X[0:rows, 0:cols] = mod(X, N), floor semantics.
 * Replaces k_mod!/mod! (src/CuModMatrix/kernel_ops/mod_ops.jl:3-27). */
int modp_mod(modp_ctx *ctx, int dtype, int64_t rows, int64_t cols, void *X, int64_t ldx, int64_t N);

/* Seeded uniform residues in [0,p): element (i,j) gets draw number
 * offset + i + j*rows of the counter-based SplitMix64 stream `seed`, mapped
 * to [0,p) by Lemire's unbiased method.  Stands in for rand(T,rows,cols,N)
 * (CuModMatrix.jl:551-556), which is unseeded in the reference. */
int modp_rand(modp_ctx *ctx, int dtype, int64_t rows, int64_t cols, void *X, int64_t ldx,
              int64_t p, uint64_t seed, uint64_t offset);

/* X = identity (rows x cols block, zeros elsewhere in the block).
 * Replaces eye(T, n, N) (CuModMatrix.jl:510-513). */
int modp_eye(modp_ctx *ctx, int dtype, int64_t rows, int64_t cols, void *X, int64_t ldx);

/* ---- GEMM ------------------------------------------------------------- */

/* C = A*B mod p; A m x k, B k x n, C m x n; entries of A and B are reduced
 * mod p on load.  Exact for 2 <= p <= 2^26 (the FP64 stripe bound
 * find_max_stripe_ops >= 1, stripe_mul.jl:13-27).
 * Replaces stripe_mul!(C,A,B) (stripe_mul.jl:175-244), used by
 * LinearAlgebra.mul!(C,A,B) (CuModMatrix.jl:767-787), Base.:* (mul_ops.jl:54-58)
 * and mat_mul_gpu_type (mat_mul_gpu_direct.jl:8-42). */
int modp_gemm(modp_ctx *ctx, int form, int dtype, int64_t m, int64_t n, int64_t k,
              const void *A, int64_t lda, const void *B, int64_t ldb,
              void *C, int64_t ldc, int64_t p);

/* modp_gemm in the limb form that allocates nothing: C = A*B mod p, with the
 * int8 digit planes of A and B held in a static device array of 64 MiB that
 * the library owns (one per device, kept until the process ends), so the
 * call makes no hipMalloc/hipFree or other allocating call.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  The
 * planes take L * (m' + n') * k' bytes (L digits of p, m' and n' rounded up to
 * the output tile, k' = min(k, 32704) rounded up to 64); larger products
 * return MODP_ERR_ARG.  Every call on a device shares that device's array:
 * calls that overlap on one device (two streams, two contexts, two threads)
 * race and must be serialised by the caller.  Kernel timing hooks are not
 * recorded for these calls. */
int modp_gemm_noalloc(modp_ctx *ctx, int dtype, int64_t m, int64_t n, int64_t k,
                      const void *A, int64_t lda, const void *B, int64_t ldb,
                      void *C, int64_t ldc, int64_t p);

/* y = A*x mod p: A m x k (column-major, leading dimension lda), x (k) and y
 * (m) contiguous device vectors; entries reduced mod p on load.  Exact for
 * 2 <= p <= 2^26.  Replaces stripe_mul!(z,A,x) (stripe_mul.jl:82-168) and
 * mul!(z,A,x) (CuModMatrix.jl:816-836). */
int modp_gemv(modp_ctx *ctx, int dtype, int64_t m, int64_t k, const void *A, int64_t lda, const void *x, void *y,
              int64_t p);

/* ---- GEMM launch configurations and the autotuner ---------------------
 * None of these has a counterpart in the reference (its only knob is the
 * stripe length of stripe_mul.jl:13-27, fixed by the modulus).  A
 * configuration is (form, output tile, K slices 1..16 run concurrently and
 * summed mod p).  MODP_FORM_AUTO products and batched plans
 * take their configuration from the context's tuning cache when the mode
 * allows, else from built-in heuristics. */
typedef struct {
    int32_t form;    /* MODP_FORM_PLAIN_F64, MODP_FORM_LIMB_I8 or MODP_FORM_RNS_I8 */
    int32_t tile_m;  /* output tile: 64x64 or 128x128; the limb form also 128x64;
                        the RNS form 128x128 or 256x256 */
    int32_t tile_n;
    int32_t ksplit;  /* 1..16 (the RNS form: 1; its grid has one layer per modulus) */
} modp_gemm_config;

int modp_autotune_mode(modp_ctx *ctx, int mode);
/* Time every valid configuration for an m x n x k product mod p on synthetic
 * operands, cache the fastest and return it. */
int modp_autotune_gemm(modp_ctx *ctx, int dtype, int64_t m, int64_t n, int64_t k, int64_t p,
                       modp_gemm_config *best, double *best_ms);
/* The configuration a MODP_FORM_AUTO product of this shape would run now;
 * *tuned = 1 when it comes from the cache. */
int modp_gemm_config_get(modp_ctx *ctx, int dtype, int64_t m, int64_t n, int64_t k, int64_t p,
                         modp_gemm_config *cfg, int *tuned);
/* modp_gemm with an explicit configuration */
int modp_gemm_ex(modp_ctx *ctx, const modp_gemm_config *cfg, int dtype, int64_t m, int64_t n, int64_t k,
                 const void *A, int64_t lda, const void *B, int64_t ldb, void *C, int64_t ldc, int64_t p);
/* The cache as text (one line per entry), to persist tuning across
 * processes.  *len receives the size including the terminating NUL; buf may
 * be NULL to query it. */
int modp_autotune_export(modp_ctx *ctx, char *buf, size_t cap, size_t *len);
int modp_autotune_import(modp_ctx *ctx, const char *text);

/* ---- elimination ------------------------------------------------------ */

/*
 * P*A*Q = L*U mod p (A m x n, not modified).  U: m x n, row echelon with unit
 * pivots; L: m x m lower triangular with L[k,k] = k-th pivot value and the
 * raw multipliers below it.  row_swaps (capacity 2*min(m,n) int64) and
 * col_swaps (capacity 4*n int64) receive the 1-based transposition stacks
 * (a,b) in order; the counts in *n_row, *n_col; the number of pivots in
 * *rank.  Pivot: largest residue of the column, lowest row on ties.  A zero
 * column is swapped with the last not-yet-swapped column and the same column
 * retried; when the remaining block is zero the stack (c, n) for c = col..n
 * is emitted.  Identical to the reference's output whenever the reference's
 * is a valid factorisation.
 * Replaces pluq_gpu_kernel(A) (src/CuModMatrix/rref_lu_pluq/pluq_kernels.jl:46-157).
 */
int modp_pluq(modp_ctx *ctx, int form, int dtype, int64_t m, int64_t n,
              const void *A, int64_t lda, void *U, int64_t ldu, void *L, int64_t ldl,
              int64_t p, int64_t *row_swaps, int64_t *n_row,
              int64_t *col_swaps, int64_t *n_col, int64_t *rank);

/*
 * PLE: A[perm,:] = L*U mod p with row pivoting only; zero columns are skipped.
 * perm: m int64, 1-based.  pivcols: min(m,n) int64, 1-based pivot columns.
 * Replaces lu_gpu_type (test/Experiments/rref_gpu_type.jl:60-103, prototype).
 */
int modp_lu(modp_ctx *ctx, int form, int dtype, int64_t m, int64_t n,
            const void *A, int64_t lda, void *U, int64_t ldu, void *L, int64_t ldl,
            int64_t p, int64_t *perm, int64_t *pivcols, int64_t *rank);

/*
 * R = canonical reduced row echelon form of A mod p (m x n); pivcols 1-based.
 * Replaces rref_gpu_type (test/Experiments/rref_gpu_type.jl:8-51, prototype;
 * README.md:8).
 */
int modp_rref(modp_ctx *ctx, int form, int dtype, int64_t m, int64_t n,
              const void *A, int64_t lda, void *R, int64_t ldr,
              int64_t p, int64_t *pivcols, int64_t *rank);

/* ---- permutations (src/CuModMatrix/rref_lu_pluq/permutations.jl) ------ */

/* Swap rows (a,b) of X for each 1-based pair of the stack, in order
 * (inverse = 0; apply_row_perm!, :73-80) or in reverse order (inverse = 1;
 * apply_row_inv_perm!, :82-89). */
int modp_apply_row_perm(modp_ctx *ctx, int dtype, int64_t rows, int64_t cols, void *X, int64_t ldx,
                        const int64_t *stack, int64_t nswaps, int inverse);
/* Same for columns (apply_col_perm!, :11-18; apply_col_inv_perm!, :20-27). */
int modp_apply_col_perm(modp_ctx *ctx, int dtype, int64_t rows, int64_t cols, void *X, int64_t ldx,
                        const int64_t *stack, int64_t nswaps, int inverse);

/* ---- multi-modular ------------------------------------------------------- */

/* Chinese remaindering of two residue matrices: Y (rows x cols) receives the
 * unique y in [0, p1*p2) with y = X1 mod p1 and y = X2 mod p2, for coprime
 * p1, p2 <= 2^26 with p1*p2 <= 2^53 (Float64) or 2^24 (Float32) so that Y is
 * exact in the storage type (a CuModArray of modulus p1*p2).  Combines the
 * per-prime results of the batched forms; the reference's nearest
 * counterpart is the two-limb composite modulus N1*N2 of KaratsubaMatrix
 * (src/KaratsubaMatrix/KaratsubaMatrix.jl:18-51). */
int modp_crt2(modp_ctx *ctx, int dtype, int64_t rows, int64_t cols, const void *X1, int64_t ld1, int64_t p1,
              const void *X2, int64_t ld2, int64_t p2, void *Y, int64_t ldy);

/* Chinese remaindering of r residue matrices (1 <= r <= 12, pairwise coprime
 * moduli p[i] in 2..2^26; X[i] a device matrix of residues mod p[i] with
 * leading dimension ldx[i], host arrays X, ldx, p) into integers, by
 * Garner's mixed-radix conversion.  Y holds `words` planes of rows x cols
 * uint32 words (leading dimension ldy, plane w at Y + w*ldy*cols), the
 * little-endian words of y in [0, P), P = prod p[i] (sign = 0), or of the
 * representative in (-P/2, P/2] in two's complement (sign = 1).  `words`
 * must cover log2(P) (+1 when signed) bits and be at most 10.  Turns the
 * per-prime results of the batched forms into one integer matrix; the
 * reference's nearest counterpart is the two-limb composite modulus of
 * KaratsubaMatrix (src/KaratsubaMatrix/KaratsubaMatrix.jl:18-51). */
int modp_crt(modp_ctx *ctx, int dtype, int64_t rows, int64_t cols, int r, const void *const *X,
             const int64_t *ldx, const int64_t *p, int sign, uint32_t *Y, int64_t ldy, int64_t words);

/* ---- inverses ------------------------------------------------------------ */

/* Inverse of a triangular matrix mod p.  uplo 0: A (m x n) is upper
 * triangular, 1: lower triangular; entries outside the triangle are ignored.
 * Ainv (n x m) receives the inverse of the leading k x k block (k = min(m,n))
 * in its leading block and zeros elsewhere, so A*Ainv = I for square and
 * wide matrices.  A zero diagonal entry gives MODP_ERR_NOT_INVERTIBLE; a tall
 * lower-triangular A gives MODP_ERR_DIM (InverseNotDefinedException).
 * Replaces upper_triangular_inverse_no_copy / lower_triangular_inverse_no_copy
 * (src/CuModMatrix/triangular/triangular_inverse_no_copy.jl:197-229,450-478). */
int modp_trtri(modp_ctx *ctx, int dtype, int uplo, int64_t m, int64_t n, const void *A, int64_t lda, void *Ainv,
               int64_t ldai, int64_t p);

/* Ainv = A^{-1} mod p for square A (n x n) through PLUQ:
 * A^{-1} = Q U^{-1} L^{-1} P.  *invertible = 0 (and Ainv untouched) when the
 * mod-p rank is below n.  Replaces inverse / is_invertible_with_inverse /
 * is_invertible (src/CuModMatrix/CuModMatrix.jl:340-502); the rank test is
 * the mod-p rank, where the reference takes a floating-point rank of U. */
int modp_inverse(modp_ctx *ctx, int form, int dtype, int64_t n, const void *A, int64_t lda, void *Ainv, int64_t ldai,
                 int64_t p, int *invertible);

/* ---- batched multi-modular forms (no counterpart in the reference; the
 *      nearest precedents are gemv_batched!, KaratsubaMatrix.jl:280, and the
 *      per-call modulus override mod_N, kernel_ops/add_ops.jl:23-30) ------- */

typedef struct {
    int64_t m, n, k, p;
    const void *A; int64_t lda;
    const void *B; int64_t ldb;
    void *C; int64_t ldc;
} modp_gemm_job;

typedef struct modp_gemm_plan modp_gemm_plan;

/* Build a plan for `count` independent products C_i = A_i*B_i mod p_i (device
 * pointers in the jobs): jobs are sorted by cost and cut into 32x32 or 64x64
 * output tiles by a threshold that the autotuner picks per batch signature
 * (under MODP_TUNE_ON_MISS a new signature is tuned by running the candidate
 * plans, which rewrites the outputs), and a persistent kernel that pulls
 * tiles from a device work queue is captured in a hipGraph. */
int modp_gemm_batched_plan(modp_ctx *ctx, int dtype, const modp_gemm_job *jobs, int64_t count,
                           modp_gemm_plan **out);
int modp_gemm_batched_run(modp_ctx *ctx, modp_gemm_plan *plan);
int modp_gemm_batched_destroy(modp_ctx *ctx, modp_gemm_plan *plan);
/* plan + run + destroy */
int modp_gemm_batched(modp_ctx *ctx, int dtype, const modp_gemm_job *jobs, int64_t count);

typedef enum { MODP_ELIM_PLUQ = 0, MODP_ELIM_LU = 1, MODP_ELIM_RREF = 2 } modp_elim_op;

typedef struct {
    int64_t m, n, p;
    const void *A; int64_t lda;  /* input, not modified */
    void *U; int64_t ldu;        /* PLUQ/LU: U; RREF: R (m x n) */
    void *L; int64_t ldl;        /* PLUQ/LU: L (m x m); RREF: unused (may be NULL) */
    int64_t *row_info;           /* device. PLUQ: row stack, 2*min(m,n) entries;
                                    LU: perm (m); RREF: pivcols (min(m,n)) */
    int64_t *col_info;           /* device. PLUQ: column stack, 4*n entries; LU: pivcols
                                    (min(m,n)); RREF: unused */
    int64_t *counts;             /* device, 3 entries: n_row, n_col, rank */
} modp_elim_job;

typedef struct modp_elim_plan modp_elim_plan;

/* Build a plan for `count` independent eliminations of small matrices
 * (m, n <= 128), each under its own prime (device pointers in the jobs):
 * the jobs are sorted by cost and their table is copied to the device once;
 * a persistent kernel, one wavefront per instance with the matrix resident in
 * LDS, pulls instances from a device work queue, and [queue reset, kernel] is
 * captured in a hipGraph.  Same pivot rule and outputs as modp_pluq /
 * modp_lu / modp_rref. */
int modp_elim_batched_plan(modp_ctx *ctx, int op, int dtype, const modp_elim_job *jobs, int64_t count,
                           modp_elim_plan **out);
/* Launch the plan's graph on the context's stream and return without
 * synchronising; the outputs are ready after modp_synchronize.  Runs of one
 * plan on one stream are ordered. */
int modp_elim_batched_run(modp_ctx *ctx, modp_elim_plan *plan);
/* Waits for the context's stream, then frees the plan. */
int modp_elim_batched_destroy(modp_ctx *ctx, modp_elim_plan *plan);
/* plan + run + wait + destroy: returns after the outputs are written. */
int modp_elim_batched(modp_ctx *ctx, int op, int dtype, const modp_elim_job *jobs, int64_t count);

/* ---- timing (for the benchmark: HIP events on the context's stream) ---- */

/* Start/stop recording of per-kernel durations for the library's hot kernels
 * (events around each launch on the context's stream). */
int modp_kernel_timing(modp_ctx *ctx, int enable);
/* Sum and count of recorded durations for kernels whose name starts with
 * `prefix`; synchronises. */
int modp_kernel_time(modp_ctx *ctx, const char *prefix, double *total_ms, int64_t *count);
int modp_kernel_timing_reset(modp_ctx *ctx);
/* events around arbitrary regions on the context's stream */
int modp_event_record(modp_ctx *ctx, int slot);
int modp_event_elapsed(modp_ctx *ctx, int slot_start, int slot_end, float *ms);
/* Measured device peak (no counterpart in the reference; for roofline
 * reports): kind 0 = dense int8 MFMA ops/s (v_mfma_i32_16x16x64_i8),
 * 1 = FP64 MFMA flop/s, 2 = HBM streaming bytes/s (1 GiB read + 1 GiB
 * written), 3 = dense int8 MFMA ops/s (v_mfma_i32_32x32x32_i8). */
int modp_probe_peak(modp_ctx *ctx, int kind, double *rate);

#ifdef __cplusplus
}
#endif

#endif /* MODP_H */
