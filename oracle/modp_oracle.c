/*
 * modp_oracle.c -- CPU oracle for mod-p GEMM and mod-p elimination.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * tests, by the package's smoke check and by the benchmark's CPU-baseline leg,
 * and by nothing else.  No product code path calls it, links it or falls back
 * to it.
 *
 * It restates, in plain C, the arithmetic of the reference package
 * GPUFiniteFieldMatrices.jl (Julia, CUDA.jl) for the hot path:
 *
 *   gemm            C = A*B mod p, exact.  The reference computes this as a
 *                   K-striped float GEMM whose partial sums stay below 2^53
 *                   (src/CuModMatrix/kernel_mul/stripe_mul.jl:175-244) and its
 *                   own CPU check is "Int64 product then mod"
 *                   (src/CuModMatrix/kernel_mul/mat_mul_bench.jl:213-216); both
 *                   equal the canonical residue of sum_k a_ik b_kj.  Here: u64
 *                   accumulation reduced every R products.
 *   mod_inv         extended Euclid, pluq_kernels.jl:11-31.
 *   pluq (src)      bit-faithful restatement of pluq_gpu_kernel,
 *                   pluq_kernels.jl:46-157 with find_pivot :179-202,
 *                   swap_cols :217-226, swap_and_mod :247-264,
 *                   swap_rows :280-289, swap_rows_and_mod :307-320,
 *                   move_and_zero_out :340-351, set_elem_kernel :366-369,
 *                   move_col_and_zero_out :385-393,
 *                   update_sub_matrix_kernel :411-440.
 *   pluq (optC)     the same pivot rule (largest residue, lowest index on
 *                   ties: Base findmax semantics at pluq_kernels.jl:188-189),
 *                   with L indexed by the pivot count and zero columns swapped
 *                   to a shrinking "last" column and the same column retried
 *                   (the intent of test/Experiments/rref_new_kernels.jl:141-147,
 *                   198-211); when the whole remaining block is zero the src
 *                   tail of column swaps (c, cols) for c = col..cols is emitted.
 *                   It equals src bit-for-bit whenever src yields a valid
 *                   factorisation.
 *   lu (PLE)        row-pivoted elimination with the same pivot rule; a zero
 *                   column is skipped (col += 1) without a column swap; returns
 *                   the row permutation vector (intent of
 *                   test/Experiments/rref_gpu_type.jl:60-103).
 *   rref            canonical reduced row echelon form by textbook
 *                   Gauss-Jordan with first-nonzero pivoting (the RREF is
 *                   unique, so the pivot rule does not matter; using a
 *                   different rule from pluq keeps the two independent).
 *   draw            the synthetic-input generator: SplitMix64 outputs taken
 *                   counter-style (output i = mix(seed + (i+1)*gamma)) then a
 *                   Lemire unbiased draw in [0,p).
 *
 * Matrices are column-major int64 with explicit leading dimension, residues
 * in [0,p).  Index stacks are returned 1-based as (a,b) pairs, matching the
 * reference's Vector{Tuple{Int,Int}}.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------ */
/* generator                                                           */
/* ------------------------------------------------------------------ */
#define SM_GAMMA 0x9E3779B97F4A7C15ULL

static inline uint64_t sm_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t oracle_draw(uint64_t seed, uint64_t index, uint64_t p)
{
    uint64_t x = sm_mix(seed + (index + 1) * SM_GAMMA);
    u128 m = (u128)x * p;
    uint64_t l = (uint64_t)m;
    if (l < p) {
        uint64_t t = (0 - p) % p;
        while (l < t) {
            x = sm_mix(x + SM_GAMMA);
            m = (u128)x * p;
            l = (uint64_t)m;
        }
    }
    return (uint64_t)(m >> 64);
}

/* Fill an rows x cols column-major block (leading dim ld) with draws whose
 * element index is offset + (column-major linear index). */
void oracle_fill(int64_t *X, int64_t rows, int64_t cols, int64_t ld,
                 uint64_t seed, uint64_t offset, uint64_t p)
{
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < cols; ++j)
        for (int64_t i = 0; i < rows; ++i)
            X[i + j * ld] = (int64_t)oracle_draw(seed, offset + (uint64_t)(i + j * rows), p);
}

/* ------------------------------------------------------------------ */
/* modular helpers                                                     */
/* ------------------------------------------------------------------ */

/* pluq_kernels.jl:11-31 */
int64_t oracle_mod_inv(int64_t a, int64_t P)
{
    int64_t inv = 0, new_inv = 1, rem = P, new_rem = a;
    while (new_rem != 0) {
        int64_t q = rem / new_rem;
        int64_t t = inv - q * new_inv; inv = new_inv; new_inv = t;
        t = rem - q * new_rem; rem = new_rem; new_rem = t;
    }
    if (inv < 0) inv += P;
    return inv;
}

/* Barrett reduction of x < 2^64 by p < 2^32: mu = floor(2^64 / p). */
typedef struct { uint64_t p, mu; } barrett_t;

static inline barrett_t barrett_make(uint64_t p)
{
    barrett_t b; b.p = p; b.mu = (uint64_t)(((u128)1 << 64) / p); return b;
}

static inline uint64_t barrett_red(barrett_t b, uint64_t x)
{
    uint64_t q = (uint64_t)(((u128)x * b.mu) >> 64);
    uint64_t r = x - q * b.p;
    while (r >= b.p) r -= b.p;
    return r;
}

/* ------------------------------------------------------------------ */
/* GEMM                                                                */
/* ------------------------------------------------------------------ */

/* C (m x n) = A (m x k) * B (k x n) mod p, all column-major, entries of A and
 * B in [0,p), p < 2^32.  u64 accumulation reduced every R products where
 * R*(p-1)^2 + p < 2^64. */
void oracle_gemm(int64_t m, int64_t n, int64_t k,
                 const int64_t *A, int64_t lda, const int64_t *B, int64_t ldb,
                 int64_t *C, int64_t ldc, int64_t p)
{
    barrett_t br = barrett_make((uint64_t)p);
    uint64_t pm1 = (uint64_t)(p - 1);
    uint64_t R = pm1 == 0 ? (uint64_t)k + 1 : (UINT64_MAX - (uint64_t)p) / (pm1 * pm1);
    if (R > (uint64_t)k) R = (uint64_t)k + 1;
    if (R < 1) R = 1;
    /* A transposed into row-major u32 so both dot-product operands are
     * contiguous over k. */
    uint32_t *At = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(m > 0 ? m : 1) * (size_t)(k > 0 ? k : 1));
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i)
        for (int64_t l = 0; l < k; ++l)
            At[i * k + l] = (uint32_t)A[i + l * lda];
    #pragma omp parallel
    {
        uint32_t *bc = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(k > 0 ? k : 1));
        #pragma omp for schedule(dynamic, 1)
        for (int64_t j = 0; j < n; ++j) {
            for (int64_t l = 0; l < k; ++l) bc[l] = (uint32_t)B[l + j * ldb];
            for (int64_t i = 0; i < m; ++i) {
                const uint32_t *a = At + i * k;
                uint64_t acc = 0;
                int64_t l = 0;
                while (l < k) {
                    int64_t e = l + (int64_t)R;
                    if (e > k) e = k;
                    uint64_t s = acc;
                    for (int64_t t = l; t < e; ++t) s += (uint64_t)a[t] * bc[t];
                    acc = barrett_red(br, s);
                    l = e;
                }
                C[i + j * ldc] = (int64_t)acc;
            }
        }
        free(bc);
    }
    free(At);
}

/* ------------------------------------------------------------------ */
/* elimination: shared state                                           */
/* ------------------------------------------------------------------ */

/* Working copy row-major: W[r*n + c]. */
typedef struct {
    int64_t m, n;
    uint64_t *W;
    barrett_t br;
} work_t;

static work_t work_load(int64_t m, int64_t n, const int64_t *A, int64_t lda, int64_t p)
{
    work_t w; w.m = m; w.n = n; w.br = barrett_make((uint64_t)p);
    w.W = (uint64_t *)calloc((size_t)(m > 0 ? m : 1) * (size_t)(n > 0 ? n : 1), sizeof(uint64_t));
    for (int64_t r = 0; r < m; ++r)
        for (int64_t c = 0; c < n; ++c)
            w.W[r * n + c] = (uint64_t)A[r + c * lda];
    return w;
}

/* argmax of W[row:m, col]: largest value, lowest row on ties.
 * ORACLE_PLANT_TIE_LAST builds a deliberately wrong variant (highest row on
 * ties) that only the tests compile, to show that the tie-break checks catch
 * it; the shipped liboracle.so never defines it. */
static void col_argmax(const work_t *w, int64_t row, int64_t col, uint64_t *v, int64_t *at)
{
    uint64_t best = 0; int64_t bi = row;
    for (int64_t r = row; r < w->m; ++r) {
        uint64_t x = w->W[r * w->n + col];
#ifdef ORACLE_PLANT_TIE_LAST
        if (x > 0 && x >= best) { best = x; bi = r; }
#else
        if (x > best) { best = x; bi = r; }
#endif
    }
    *v = best; *at = bi;
}

static void swap_cols(work_t *w, int64_t c1, int64_t c2)
{
    if (c1 == c2) return;
    for (int64_t r = 0; r < w->m; ++r) {
        uint64_t t = w->W[r * w->n + c1];
        w->W[r * w->n + c1] = w->W[r * w->n + c2];
        w->W[r * w->n + c2] = t;
    }
}

/* pluq_kernels.jl:307-320 with row1 = prow, row2 = row over all columns:
 * old pivot row moved to `row` and multiplied by inv, old `row` moved to
 * prow. */
static void swap_rows_and_scale(work_t *w, int64_t row, int64_t prow, uint64_t inv)
{
    uint64_t *a = w->W + row * w->n, *b = w->W + prow * w->n;
    for (int64_t c = 0; c < w->n; ++c) {
        uint64_t t = b[c];
        b[c] = a[c];
        a[c] = barrett_red(w->br, t * inv);
    }
}

/* rank-1 update pluq_kernels.jl:411-440: for r > row, c > col,
 * A[r,c] = (A[r,c] + (p - mult_r) * A[row,c]) mod p, with mult_r taken from
 * multcol (the L column).  Only columns < ncols are touched. */
static void rank1_update(work_t *w, int64_t row, int64_t col, const uint64_t *mult)
{
    const uint64_t p = w->br.p;
    const uint64_t *prow = w->W + row * w->n;
    #pragma omp parallel for schedule(static) if ((w->m - row) * (w->n - col) > 65536)
    for (int64_t r = row + 1; r < w->m; ++r) {
        uint64_t f = p - mult[r];
        if (f == p) continue;          /* (p - 0) * u == 0 mod p */
        uint64_t *a = w->W + r * w->n;
        for (int64_t c = col + 1; c < w->n; ++c)
            a[c] = barrett_red(w->br, a[c] + f * prow[c]);
    }
}

static void work_store(const work_t *w, int64_t *U, int64_t ldu)
{
    for (int64_t c = 0; c < w->n; ++c)
        for (int64_t r = 0; r < w->m; ++r)
            U[r + c * ldu] = (int64_t)w->W[r * w->n + c];
}

/* ------------------------------------------------------------------ */
/* PLUQ                                                                */
/* ------------------------------------------------------------------ */

enum { ORACLE_PLUQ_SRC = 0, ORACLE_PLUQ_OPTC = 1 };

/*
 * A: m x n column-major (lda).  U: m x n (ldu).  L: m x lcols (ldl) where
 * lcols = max(m, n): src mode indexes L by the column of the pivot
 * (pluq_kernels.jl:343,347,430), which can exceed m for wide inputs; optC
 * only ever writes the first m columns.
 * row_swaps / col_swaps: 2*cap int64 each (cap >= 2*n + m); pairs 1-based.
 * Returns 0, or -1 if a stack would overflow cap.
 */
int oracle_pluq(int mode, int64_t m, int64_t n, const int64_t *A, int64_t lda, int64_t p,
                int64_t *U, int64_t ldu, int64_t *L, int64_t ldl,
                int64_t *row_swaps, int64_t *n_row, int64_t *col_swaps, int64_t *n_col,
                int64_t cap, int64_t *rank)
{
    work_t w = work_load(m, n, A, lda, p);
    int64_t lcols = m > n ? m : n;
    uint64_t *Lw = (uint64_t *)calloc((size_t)(m > 0 ? m : 1) * (size_t)(lcols > 0 ? lcols : 1), sizeof(uint64_t));
#define LW(r, c) Lw[(r) * lcols + (c)]
    uint64_t *mult = (uint64_t *)calloc((size_t)(m > 0 ? m : 1), sizeof(uint64_t));
    int64_t nr = 0, nc = 0, row = 0, col = 0, last = n - 1;
    int rc = 0;

    while (row < m && col < n) {
        uint64_t v = 0; int64_t at = row;
        int stop = 0;
        for (;;) {
            col_argmax(&w, row, col, &v, &at);
            if (v > 0) break;
            if (mode == ORACLE_PLUQ_SRC) {
                /* find_pivot :191-198 then the caller's col += 1 :100-107 */
                if (nc >= cap) { rc = -1; goto done; }
                swap_cols(&w, col, n - 1);
                col_swaps[2 * nc] = col + 1; col_swaps[2 * nc + 1] = n; ++nc;
                ++col;
                if (col >= n) { stop = 1; break; }
            } else {
                /* optC: is the whole block W[row:m, col:n] zero? */
                int allzero = 1;
                for (int64_t r = row; r < m && allzero; ++r)
                    for (int64_t c = col; c < n; ++c)
                        if (w.W[r * n + c]) { allzero = 0; break; }
                if (allzero) {
                    for (int64_t c = col; c < n; ++c) {
                        if (nc >= cap) { rc = -1; goto done; }
                        swap_cols(&w, c, n - 1);
                        col_swaps[2 * nc] = c + 1; col_swaps[2 * nc + 1] = n; ++nc;
                    }
                    stop = 1; break;
                }
                if (nc >= cap) { rc = -1; goto done; }
                swap_cols(&w, col, last);
                col_swaps[2 * nc] = col + 1; col_swaps[2 * nc + 1] = last + 1; ++nc;
                --last;
            }
        }
        if (stop) break;

        uint64_t inv = (uint64_t)oracle_mod_inv((int64_t)v, p);
        /* swap_and_mod :247-264 */
        swap_rows_and_scale(&w, row, at, inv);
        if (at != row) {
            /* swap_rows on d_L covers columns 1..32*cld(rows,32) (:254, :282) */
            int64_t sw = ((m + 31) / 32) * 32;
            if (sw > lcols) sw = lcols;
            for (int64_t c = 0; c < sw; ++c) { uint64_t t = LW(row, c); LW(row, c) = LW(at, c); LW(at, c) = t; }
            if (nr >= cap) { rc = -1; goto done; }
            row_swaps[2 * nr] = row + 1; row_swaps[2 * nr + 1] = at + 1; ++nr;
        }
        /* move_and_zero_out :340-351 */
        int64_t lc = (mode == ORACLE_PLUQ_SRC) ? col : row;
        LW(row, lc) = v;
        for (int64_t r = row + 1; r < m; ++r) {
            LW(r, lc) = w.W[r * n + col];
            mult[r] = w.W[r * n + col];
            w.W[r * n + col] = 0;
        }
        /* update_sub_matrix_kernel :411-440 */
        rank1_update(&w, row, col, mult);
        ++row; ++col;
    }
done:
    work_store(&w, U, ldu);
    for (int64_t c = 0; c < lcols; ++c)
        for (int64_t r = 0; r < m; ++r)
            L[r + c * ldl] = (int64_t)LW(r, c);
#undef LW
    *n_row = nr; *n_col = nc; *rank = row;
    free(mult); free(Lw); free(w.W);
    return rc;
}

/* ------------------------------------------------------------------ */
/* LU (PLE): row pivoting only, zero columns skipped                   */
/* ------------------------------------------------------------------ */

/* perm: m entries, 1-based, A[perm,:] = L*U.  pivcols: 1-based pivot column
 * of each pivot (rank entries). */
int oracle_lu(int64_t m, int64_t n, const int64_t *A, int64_t lda, int64_t p,
              int64_t *U, int64_t ldu, int64_t *L, int64_t ldl,
              int64_t *perm, int64_t *pivcols, int64_t *rank)
{
    work_t w = work_load(m, n, A, lda, p);
    uint64_t *Lw = (uint64_t *)calloc((size_t)(m > 0 ? m : 1) * (size_t)(m > 0 ? m : 1), sizeof(uint64_t));
    uint64_t *mult = (uint64_t *)calloc((size_t)(m > 0 ? m : 1), sizeof(uint64_t));
    for (int64_t i = 0; i < m; ++i) perm[i] = i + 1;
    int64_t row = 0, col = 0;
    while (row < m && col < n) {
        uint64_t v; int64_t at;
        col_argmax(&w, row, col, &v, &at);
        if (v == 0) { ++col; continue; }
        uint64_t inv = (uint64_t)oracle_mod_inv((int64_t)v, p);
        swap_rows_and_scale(&w, row, at, inv);
        if (at != row) {
            for (int64_t c = 0; c < m; ++c) { uint64_t t = Lw[row * m + c]; Lw[row * m + c] = Lw[at * m + c]; Lw[at * m + c] = t; }
            int64_t t = perm[row]; perm[row] = perm[at]; perm[at] = t;
        }
        Lw[row * m + row] = v;
        for (int64_t r = row + 1; r < m; ++r) {
            Lw[r * m + row] = w.W[r * n + col];
            mult[r] = w.W[r * n + col];
            w.W[r * n + col] = 0;
        }
        rank1_update(&w, row, col, mult);
        pivcols[row] = col + 1;
        ++row; ++col;
    }
    work_store(&w, U, ldu);
    for (int64_t c = 0; c < m; ++c)
        for (int64_t r = 0; r < m; ++r)
            L[r + c * ldl] = (int64_t)Lw[r * m + c];
    *rank = row;
    free(mult); free(Lw); free(w.W);
    return 0;
}

/* ------------------------------------------------------------------ */
/* RREF: textbook Gauss-Jordan, first nonzero pivot                    */
/* ------------------------------------------------------------------ */

int oracle_rref(int64_t m, int64_t n, const int64_t *A, int64_t lda, int64_t p,
                int64_t *R, int64_t ldr, int64_t *pivcols, int64_t *rank)
{
    work_t w = work_load(m, n, A, lda, p);
    const uint64_t P = (uint64_t)p;
    int64_t row = 0;
    for (int64_t col = 0; col < n && row < m; ++col) {
        int64_t at = -1;
        for (int64_t r = row; r < m; ++r) if (w.W[r * n + col]) { at = r; break; }
        if (at < 0) continue;
        if (at != row)
            for (int64_t c = 0; c < n; ++c) { uint64_t t = w.W[row * n + c]; w.W[row * n + c] = w.W[at * n + c]; w.W[at * n + c] = t; }
        uint64_t inv = (uint64_t)oracle_mod_inv((int64_t)w.W[row * n + col], p);
        uint64_t *pr = w.W + row * n;
        for (int64_t c = col; c < n; ++c) pr[c] = barrett_red(w.br, pr[c] * inv);
        #pragma omp parallel for schedule(static) if (m * (n - col) > 65536)
        for (int64_t r = 0; r < m; ++r) {
            if (r == row) continue;
            uint64_t *a = w.W + r * n;
            uint64_t f = a[col];
            if (!f) continue;
            f = P - f;
            for (int64_t c = col; c < n; ++c) a[c] = barrett_red(w.br, a[c] + f * pr[c]);
        }
        pivcols[row] = col + 1;
        ++row;
    }
    work_store(&w, R, ldr);
    *rank = row;
    free(w.W);
    return 0;
}

/* ------------------------------------------------------------------ */
/* primes                                                              */
/* ------------------------------------------------------------------ */

static uint64_t powmod(uint64_t a, uint64_t e, uint64_t m)
{
    u128 r = 1, b = a % m;
    while (e) { if (e & 1) r = (r * b) % m; b = (b * b) % m; e >>= 1; }
    return (uint64_t)r;
}

/* deterministic Miller-Rabin for n < 3.3e24 with the first 12 prime bases */
int oracle_is_prime(uint64_t n)
{
    static const uint64_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (int i = 0; i < 12; ++i) { if (n == bases[i]) return 1; if (n % bases[i] == 0) return 0; }
    uint64_t d = n - 1; int s = 0;
    while (!(d & 1)) { d >>= 1; ++s; }
    for (int i = 0; i < 12; ++i) {
        uint64_t x = powmod(bases[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int comp = 1;
        for (int r = 1; r < s; ++r) {
            x = (uint64_t)(((u128)x * x) % n);
            if (x == n - 1) { comp = 0; break; }
        }
        if (comp) return 0;
    }
    return 1;
}
