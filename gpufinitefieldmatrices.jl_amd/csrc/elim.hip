// elim.hip -- blocked right-looking mod-p elimination on gfx950.
//
// One driver serves pluq (column pivoting of zero columns), lu (PLE: zero
// columns skipped) and rref (PLE followed by blocked back-elimination).  The
// matrix is copied into an int32 working array W (column-major) that keeps L
// and U in place: the pivot of pivot number k sits at row k, column pivcol[k]
// (== k for pluq), the multipliers below it are L's column k, and the
// normalised pivot row is U's row k.
//
// Structure per outer block of NBO columns (reference: the unblocked loop of
// src/CuModMatrix/rref_lu_pluq/pluq_kernels.jl:90-142, one pivot per host
// iteration with a host round-trip each; here one launch factors a whole
// panel):
//   for each inner panel of w columns (w*rows-per-thread <= 64 registers):
//     k_panel    one 1024-thread workgroup holds the panel in registers; per
//                column: wave64 argmax by __shfl_xor with lowest-row
//                tie-break, cross-wave through LDS, modular inverse of the
//                pivot by extended Euclid, row swap, normalisation and the
//                rank-1 update of the remaining panel columns.
//     k_finish   one wave per column outside the panel: the panel's row
//                swaps (as a precomputed gather map), then for columns right
//                of the panel the triangular solve with the panel's L11
//                (lanes hold the column, shuffles broadcast the solved rows),
//                writing U's rows into the outer-block buffer Ubig.
//     G1, G2     GEMM updates (EPI_SUB) of the rest of the outer block (all
//                rows) and of the rows of the far columns that are still in
//                this outer block.
//   after the outer block: one GEMM with K = NBO for the far trailing block.
#include "common.h"

namespace {

constexpr int NBO = 128;  // outer block width (K of the trailing GEMM)

struct PanelCtl {
    int npiv;    // pivots found
    int status;  // 0 all columns consumed, 1 zero column (pluq stop),
                 // 2 rows exhausted, 3 pivot not invertible
    int ncols;   // panel columns consumed
    int nmap;    // entries of the row gather map
};

template <int RPT, int NBW>
__global__ __launch_bounds__(1024) void k_panel(int32_t *__restrict__ W, int64_t ldw, int m, int kr0, int c0,
                                                int w, ModP mp, int ple, int32_t *__restrict__ ipiv,
                                                int32_t *__restrict__ pivval, int32_t *__restrict__ pivinv,
                                                int32_t *__restrict__ pivcol, int32_t *__restrict__ colmap,
                                                int32_t *__restrict__ map_dst, int32_t *__restrict__ map_src,
                                                PanelCtl *__restrict__ ctl)
{
    __shared__ uint32_t s_val[2][16];
    __shared__ int s_idx[2][16];
    __shared__ uint32_t s_prow[NBW], s_crow[NBW], s_nrow[NBW];
    __shared__ uint32_t s_inv;
    __shared__ int s_ipiv[NBW];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rows = m - kr0;
    uint32_t v[RPT][NBW];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int i = tid + 1024 * q;
#pragma unroll
        for (int j = 0; j < NBW; ++j)
            v[q][j] = (i < rows && j < w) ? (uint32_t)W[(int64_t)(kr0 + i) + (int64_t)(c0 + j) * ldw] : 0u;
    }

    int pc = 0, status = 0, consumed = w, par = 0;
    bool active = true;
    // fully unrolled over the panel's columns so that v[][] stays in registers
#pragma unroll
    for (int jc = 0; jc < NBW; ++jc) {
        if (!active || jc >= w) continue;
        if (pc >= rows) {
            status = 2;
            consumed = jc;
            active = false;
            continue;
        }
        // candidate of this thread: largest value, lowest row
        uint32_t bv = 0;
        int bi = 0x7fffffff;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int i = tid + 1024 * q;
            if (i >= pc && i < rows) {
                uint32_t x = v[q][jc];
                if (x > bv) { bv = x; bi = i; }
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            uint32_t ov = __shfl_xor(bv, off);
            int oi = __shfl_xor(bi, off);
            if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
        }
        if (lane == 0) { s_val[par][wave] = bv; s_idx[par][wave] = bi; }
        __syncthreads();
        uint32_t V = s_val[par][0];
        int I = s_idx[par][0];
#pragma unroll
        for (int x = 1; x < 16; ++x) {
            uint32_t ov = s_val[par][x];
            int oi = s_idx[par][x];
            if (ov > V || (ov == V && oi < I)) { V = ov; I = oi; }
        }
        par ^= 1;
        if (V == 0) {
            if (ple) {
                if (tid == 0) colmap[jc] = -1;
            } else {
                status = 1;
                consumed = jc;
                active = false;
            }
            continue;
        }
        const int tI = I & 1023, qI = I >> 10, tP = pc & 1023, qP = pc >> 10;
        if (tid == tI) {
            const uint32_t inv = modinv(V, mp);
            s_inv = inv;
#pragma unroll
            for (int q = 0; q < RPT; ++q)
                if (q == qI) {
#pragma unroll
                    for (int j = 0; j < NBW; ++j) {
                        s_prow[j] = v[q][j];
                        if (j > jc) s_nrow[j] = mulmod(v[q][j], inv, mp);
                    }
                }
        }
        if (tid == tP) {
#pragma unroll
            for (int q = 0; q < RPT; ++q)
                if (q == qP) {
#pragma unroll
                    for (int j = 0; j < NBW; ++j) s_crow[j] = v[q][j];
                }
        }
        __syncthreads();
        const uint32_t inv = s_inv;
        if (inv == 0) {
            status = 3;
            consumed = jc;
            active = false;
            continue;
        }
        if (tid == tI) {
#pragma unroll
            for (int q = 0; q < RPT; ++q)
                if (q == qI) {
#pragma unroll
                    for (int j = 0; j < NBW; ++j) v[q][j] = s_crow[j];
                }
        }
        if (tid == tP) {
#pragma unroll
            for (int q = 0; q < RPT; ++q)
                if (q == qP) {
#pragma unroll
                    for (int j = 0; j < NBW; ++j) v[q][j] = j < jc ? s_prow[j] : (j == jc ? 1u : s_nrow[j]);
                }
        }
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int i = tid + 1024 * q;
            if (i > pc && i < rows) {
                const uint32_t l = v[q][jc];
                if (l) {
#pragma unroll
                    for (int j = 0; j < NBW; ++j)
                        if (j > jc) v[q][j] = submod(v[q][j], mulmod(l, s_nrow[j], mp), mp);
                }
            }
        }
        if (tid == 0) {
            ipiv[kr0 + pc] = kr0 + I;
            pivval[kr0 + pc] = (int32_t)V;
            pivinv[kr0 + pc] = (int32_t)inv;
            pivcol[kr0 + pc] = c0 + jc;
            colmap[jc] = pc;
            s_ipiv[pc] = kr0 + I;
        }
        ++pc;
    }

#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int i = tid + 1024 * q;
        if (i < rows) {
#pragma unroll
            for (int j = 0; j < NBW; ++j)
                if (j < w) W[(int64_t)(kr0 + i) + (int64_t)(c0 + j) * ldw] = (int32_t)v[q][j];
        }
    }

    if (tid == 0) {
        for (int j = consumed; j < w; ++j) colmap[j] = -1;
        // gather map of the row swaps: rows kr0..kr0+pc-1 first, in order,
        // then every other row whose content changed
        int rowv[2 * NBW], orig[2 * NBW], cnt = 0;
        for (int i = 0; i < pc; ++i) {
            const int a = kr0 + i, b = s_ipiv[i];
            if (a == b) continue;
            int ia = -1, ib = -1;
            for (int e = 0; e < cnt; ++e) {
                if (rowv[e] == a) ia = e;
                if (rowv[e] == b) ib = e;
            }
            if (ia < 0) { rowv[cnt] = a; orig[cnt] = a; ia = cnt++; }
            if (ib < 0) { rowv[cnt] = b; orig[cnt] = b; ib = cnt++; }
            int t = orig[ia]; orig[ia] = orig[ib]; orig[ib] = t;
        }
        int n = 0;
        for (int i = 0; i < pc; ++i) {
            int src = kr0 + i;
            for (int e = 0; e < cnt; ++e)
                if (rowv[e] == kr0 + i) src = orig[e];
            map_dst[n] = kr0 + i;
            map_src[n] = src;
            ++n;
        }
        for (int e = 0; e < cnt; ++e)
            if (rowv[e] >= kr0 + pc && orig[e] != rowv[e]) {
                map_dst[n] = rowv[e];
                map_src[n] = orig[e];
                ++n;
            }
        ctl->npiv = pc;
        ctl->status = status;
        ctl->ncols = consumed;
        ctl->nmap = n;
    }
}

// One wave per column of the outer block [ca, cb) outside the panel
// [c0, c0+w): apply the panel's row swaps (a gather map whose first npiv
// entries are the pivot rows in order); for columns right of the panel also
// solve with the panel's L11 and write the U rows into Ubig (rows indexed by
// pivot column - KC, columns by c - KC).
__global__ __launch_bounds__(256) void k_finish(int32_t *__restrict__ W, int64_t ldw, int64_t ca, int64_t cb,
                                                int kr0, int c0, int w, int KC, const PanelCtl *__restrict__ ctl,
                                                const int32_t *__restrict__ pivinv,
                                                const int32_t *__restrict__ pivcol,
                                                const int32_t *__restrict__ colmap,
                                                const int32_t *__restrict__ map_dst,
                                                const int32_t *__restrict__ map_src, int32_t *__restrict__ Ubig,
                                                int64_t ldub, ModP mp)
{
    __shared__ uint32_t Ls[64][65];
    __shared__ uint32_t invs[64];
    __shared__ int pcols[64];
    const int npiv = ctl->npiv, nmap = ctl->nmap;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int idx = tid; idx < npiv * npiv; idx += 256) {
        const int i = idx / npiv, j = idx % npiv;
        Ls[i][j] = j < i ? (uint32_t)W[(int64_t)(kr0 + i) + (int64_t)pivcol[kr0 + j] * ldw] : 0u;
    }
    if (tid < npiv) {
        invs[tid] = (uint32_t)pivinv[kr0 + tid];
        pcols[tid] = pivcol[kr0 + tid];
    }
    __syncthreads();
    const int64_t ncols = (cb - ca) - w;
    const int d0 = lane < nmap ? map_dst[lane] : 0, s0 = lane < nmap ? map_src[lane] : 0;
    const bool zero_row = lane < w && colmap[lane] < 0;
    for (int64_t ci = (int64_t)blockIdx.x * 4 + wave; ci < ncols; ci += (int64_t)gridDim.x * 4) {
        int64_t c = ca + ci;
        if (c >= c0) c += w;
        int32_t *col = W + c * ldw;
        uint32_t v0 = 0;
        if (lane < nmap) v0 = (uint32_t)col[s0];
        if (c < c0) {
            if (lane < nmap) col[d0] = (int32_t)v0;
            continue;
        }
        if (lane >= npiv && lane < nmap) col[d0] = (int32_t)v0;
        uint32_t x = v0;
        for (int j = 0; j < npiv; ++j) {
            if (lane == j) x = mulmod(x, invs[j], mp);
            const uint32_t xj = (uint32_t)__shfl((int)x, j);
            if (lane > j && lane < npiv) x = submod(x, mulmod(Ls[lane][j], xj, mp), mp);
        }
        int32_t *ucol = Ubig + (c - KC) * ldub;
        if (lane < npiv) {
            col[kr0 + lane] = (int32_t)x;
            ucol[pcols[lane] - KC] = (int32_t)x;
        }
        if (zero_row) ucol[c0 - KC + lane] = 0;
    }
}

constexpr int MAXPAN = 32;  // inner panels per outer block (NBO / smallest panel width)

// End of an outer block with P pivots (rows [KR, KR+P)), whose inner panels
// left their row gather maps in bmap_*: one wave per column outside the block
// [KC, FAR).  Columns left of the block get the swaps; far columns get the
// swaps, then the solve with the block's L11 (P x P, in LDS; lanes hold rows
// i and i+64), and their U rows go to Ubig (rows 0..consumed-1 indexed by
// block column, zero for block columns without a pivot).
__global__ __launch_bounds__(1024) void k_block_finish(int32_t *__restrict__ W, int64_t ldw, int64_t n, int KR,
                                                       int KC, int FAR, int P, int consumed, int npanels,
                                                       const int32_t *__restrict__ bmap_dst,
                                                       const int32_t *__restrict__ bmap_src,
                                                       const PanelCtl *__restrict__ bctl,
                                                       const int32_t *__restrict__ pivinv,
                                                       const int32_t *__restrict__ pivcol,
                                                       int32_t *__restrict__ Ubig, int64_t ldub, ModP mp)
{
    __shared__ uint32_t Ls[128][129];
    __shared__ uint32_t invs[128];
    __shared__ int ucolmap[128];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int idx = tid; idx < P * P; idx += 1024) {
        const int i = idx / P, j = idx % P;
        Ls[i][j] = j < i ? (uint32_t)W[(int64_t)(KR + i) + (int64_t)pivcol[KR + j] * ldw] : 0u;
    }
    if (tid < 128) ucolmap[tid] = -1;
    __syncthreads();
    if (tid < P) {
        invs[tid] = (uint32_t)pivinv[KR + tid];
        ucolmap[pivcol[KR + tid] - KC] = tid;
    }
    __syncthreads();
    const int64_t nout = n - (FAR - KC);
    const int um0 = lane < consumed ? ucolmap[lane] : -1;
    const int um1 = lane + 64 < consumed ? ucolmap[lane + 64] : -1;
    for (int64_t ci = (int64_t)blockIdx.x * 16 + wave; ci < nout; ci += (int64_t)gridDim.x * 16) {
        int64_t c = ci < KC ? ci : ci + (FAR - KC);
        int32_t *col = W + c * ldw;
        for (int t = 0; t < npanels; ++t) {
            const int nm = bctl[t].nmap;
            const int32_t *md = bmap_dst + t * 64, *ms = bmap_src + t * 64;
            uint32_t v = 0;
            int d = 0;
            if (lane < nm) { v = (uint32_t)col[ms[lane]]; d = md[lane]; }
            if (lane < nm) col[d] = (int32_t)v;
            __threadfence_block();
        }
        if (c < KC) continue;
        uint32_t x0 = lane < P ? (uint32_t)col[KR + lane] : 0u;
        uint32_t x1 = lane + 64 < P ? (uint32_t)col[KR + lane + 64] : 0u;
        for (int j = 0; j < P; ++j) {
            uint32_t xj;
            if (j < 64) {
                if (lane == j) x0 = mulmod(x0, invs[j], mp);
                xj = (uint32_t)__shfl((int)x0, j);
                if (lane > j && lane < P) x0 = submod(x0, mulmod(Ls[lane][j], xj, mp), mp);
            } else {
                if (lane == j - 64) x1 = mulmod(x1, invs[j], mp);
                xj = (uint32_t)__shfl((int)x1, j - 64);
            }
            if (lane + 64 > j && lane + 64 < P) x1 = submod(x1, mulmod(Ls[lane + 64][j], xj, mp), mp);
        }
        if (lane < P) col[KR + lane] = (int32_t)x0;
        if (lane + 64 < P) col[KR + lane + 64] = (int32_t)x1;
        // Ubig rows 0..consumed-1 of this column: pivot i's value or zero
        int32_t *ucol = Ubig + (c - KC) * ldub;
        const int a0 = um0 < 0 ? 0 : um0, a1 = um1 < 0 ? 0 : um1;
        const uint32_t y00 = (uint32_t)__shfl((int)x0, a0 & 63), y01 = (uint32_t)__shfl((int)x1, a0 & 63);
        const uint32_t y10 = (uint32_t)__shfl((int)x0, a1 & 63), y11 = (uint32_t)__shfl((int)x1, a1 & 63);
        if (lane < consumed) ucol[lane] = um0 < 0 ? 0 : (int32_t)(a0 < 64 ? y00 : y01);
        if (lane + 64 < consumed) ucol[lane + 64] = um1 < 0 ? 0 : (int32_t)(a1 < 64 ? y10 : y11);
    }
}

// any nonzero in W[r0:m, c0:n] -> *flag = 1
__global__ void k_any_nonzero(const int32_t *__restrict__ W, int64_t ldw, int64_t r0, int64_t m, int64_t c0,
                              int64_t n, int *flag)
{
    const int64_t rows = m - r0, total = rows * (n - c0);
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = r0 + e % rows, j = c0 + e / rows;
        if (W[i + j * ldw] != 0) *flag = 1;
    }
}

template <typename TU>
__global__ void k_extract_u(int64_t m, int64_t n, const int32_t *__restrict__ W, int64_t ldw, int rank,
                            const int32_t *__restrict__ pivcol, TU *__restrict__ U, int64_t ldu)
{
    const int64_t total = m * n;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e % m, c = e / m;
        uint32_t v = 0;
        if (r < rank && c >= pivcol[r]) v = (uint32_t)W[r + c * ldw];
        U[r + c * ldu] = from_res<TU>(v);
    }
}

template <typename TL>
__global__ void k_extract_l(int64_t m, const int32_t *__restrict__ W, int64_t ldw, int rank,
                            const int32_t *__restrict__ pivcol, const int32_t *__restrict__ pivval,
                            TL *__restrict__ L, int64_t ldl)
{
    const int64_t total = m * m;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e % m, j = e / m;
        uint32_t v = 0;
        if (j < rank) {
            if (r == j) v = (uint32_t)pivval[j];
            else if (r > j) v = (uint32_t)W[r + (int64_t)pivcol[j] * ldw];
        }
        L[r + j * ldl] = from_res<TL>(v);
    }
}

// clear the multipliers kept below the pivots: W[r, pivcol[j]] = 0 for r > j
__global__ void k_zero_lower(int32_t *__restrict__ W, int64_t ldw, int64_t m, int rank,
                             const int32_t *__restrict__ pivcol)
{
    const int64_t total = m * rank;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e % m, j = e / m;
        if (r > j) W[r + (int64_t)pivcol[j] * ldw] = 0;
    }
}

// G[i, j] = W[r0 + i, pivcol[p0 + j]] for i < rows, j < nb
__global__ void k_gather_cols(const int32_t *__restrict__ W, int64_t ldw, int64_t r0, int64_t rows,
                              const int32_t *__restrict__ pivcol, int p0, int nb, int32_t *__restrict__ G,
                              int64_t ldg)
{
    const int64_t total = rows * nb;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % rows, j = e / rows;
        G[i + j * ldg] = W[r0 + i + (int64_t)pivcol[p0 + j] * ldw];
    }
}

// Back substitution with a unit upper triangular T (nb x nb, gathered into
// G) on rows [i0, i0+nb) of columns [cs, n): one wave per column.
__global__ __launch_bounds__(256) void k_trsm_upper_unit(int32_t *__restrict__ W, int64_t ldw, int64_t n, int i0,
                                                         int nb, int64_t cs, const int32_t *__restrict__ G,
                                                         int64_t ldg, ModP mp)
{
    __shared__ uint32_t Ts[64][65];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int idx = tid; idx < nb * nb; idx += 256) {
        const int i = idx % nb, j = idx / nb;
        Ts[i][j] = (uint32_t)G[i + j * ldg];
    }
    __syncthreads();
    for (int64_t c = cs + (int64_t)blockIdx.x * 4 + wave; c < n; c += (int64_t)gridDim.x * 4) {
        int32_t *col = W + c * ldw + i0;
        uint32_t x = lane < nb ? (uint32_t)col[lane] : 0u;
        for (int j = nb - 1; j > 0; --j) {
            const uint32_t xj = (uint32_t)__shfl((int)x, j);
            if (lane < j) x = submod(x, mulmod(Ts[lane][j], xj, mp), mp);
        }
        if (lane < nb) col[lane] = (int32_t)x;
    }
}

template <int RPT, int NBW>
int launch_panel(modp_ctx *ctx, int32_t *W, int64_t ldw, int m, int kr0, int c0, int w, const ModP &mp, int ple,
                 int32_t *ipiv, int32_t *pivval, int32_t *pivinv, int32_t *pivcol, int32_t *colmap,
                 int32_t *map_dst, int32_t *map_src, PanelCtl *ctl)
{
    TimedLaunch tl(ctx, "elim_panel");
    k_panel<RPT, NBW><<<1, 1024, 0, ctx->stream>>>(W, ldw, m, kr0, c0, w, mp, ple, ipiv, pivval, pivinv, pivcol,
                                                   colmap, map_dst, map_src, ctl);
    return hip_check(ctx, hipGetLastError(), "k_panel launch");
}

unsigned grid_for(int64_t total, int64_t cap = 16384)
{
    int64_t b = (total + 255) / 256;
    if (b > cap) b = cap;
    if (b < 1) b = 1;
    return (unsigned)b;
}

int64_t rup(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

}  // namespace

int elim_run(modp_ctx *ctx, int mode, int form, int kind, int64_t m, int64_t n, const void *A, int64_t lda,
             void *U, int64_t ldu, void *L, int64_t ldl, int64_t p, ElimResult &res)
{
    const ModP mp = make_modp(p);
    const bool ple = mode != ELIM_PLUQ;
    const int64_t mn = m < n ? m : n;
    res = ElimResult();
    if (m <= 0 || n <= 0) {
        if (mode == ELIM_LU)
            for (int64_t i = 0; i < m; ++i) res.perm.push_back(i + 1);
        if (U && m > 0 && n > 0) return MODP_OK;
        return MODP_OK;
    }
    if (m > 16384)
        return set_error(ctx, MODP_ERR_DIM, "elimination supports at most 16384 rows (got %lld)", (long long)m);

    // panel shape: rows per thread x panel width <= 64 registers
    int RPT = 1;
    while ((int64_t)RPT * 1024 < m) RPT *= 2;
    const int NBW = RPT <= 2 ? 32 : 64 / RPT;

    int st = MODP_OK;
    const int64_t ldw = rup(m, 64);
    int32_t *W = (int32_t *)ws_get(ctx, WS_WORK, sizeof(int32_t) * (size_t)ldw * (size_t)n, &st);
    if (!W) return st;
    // pivot arrays, per-panel control blocks and row gather maps of one outer block
    const size_t npv = (size_t)mn + 64;
    const size_t piv_bytes = sizeof(int32_t) * (4 * npv + 64 + 2 * MAXPAN * 64) + sizeof(PanelCtl) * MAXPAN + 64;
    int32_t *piv = (int32_t *)ws_get(ctx, WS_PIV, piv_bytes, &st);
    if (!piv) return st;
    int32_t *ipiv = piv, *pivval = piv + npv, *pivinv = piv + 2 * npv, *pivcol = piv + 3 * npv;
    int32_t *colmap = piv + 4 * npv;
    int32_t *bmap_dst = colmap + 64, *bmap_src = bmap_dst + MAXPAN * 64;
    PanelCtl *bctl = (PanelCtl *)(bmap_src + MAXPAN * 64);
    int *flag = (int *)(bctl + MAXPAN);
    int32_t *Ubig = (int32_t *)ws_get(ctx, WS_UBUF, sizeof(int32_t) * (size_t)NBO * (size_t)(n + 1), &st);
    if (!Ubig) return st;
    PanelCtl *hctl = (PanelCtl *)pinned_get(ctx, 64, &st);
    if (!hctl) return st;

    MODP_TRY(util_convert(ctx, kind, m, n, A, lda, EK_I32, W, ldw, mp));

    auto gemm_sub = [&](int64_t r0, int64_t rows, int64_t ca, int64_t cols, int64_t kcol, int64_t K,
                        const int32_t *Ub, int use_form) -> int {
        if (rows <= 0 || cols <= 0 || K <= 0) return MODP_OK;
        GemmArgs g;
        g.m = rows; g.n = cols; g.k = K;
        g.A = W + r0 + kcol * ldw; g.lda = ldw; g.ka = EK_I32;
        g.B = Ub; g.ldb = NBO; g.kb = EK_I32;
        g.C = W + r0 + ca * ldw; g.ldc = ldw; g.kc = EK_I32;
        g.epi = EPI_SUB;
        g.mp = mp;
        if (use_form == MODP_FORM_LIMB_I8) return gemm_limb_i8(ctx, g);
        return gemm_plain_f64(ctx, g);
    };
    const int big_form = form == MODP_FORM_AUTO ? choose_form(m, n, NBO, p) : form;

    std::vector<int64_t> col_stack;
    int64_t kr = 0, kc = 0, last = n - 1;
    bool done = false;
    while (!done && kr < m && kc < n) {
        // ---- one outer block of columns [KC, FAR) ----
        const int64_t KR = kr, KC = kc;
        const int64_t colend = ple ? n : last + 1;  // columns usable as pivot columns
        int64_t nbo = colend - KC;
        if (nbo > NBO) nbo = NBO;
        if (nbo < 0) nbo = 0;  // pluq: only moved zero columns remain
        const int64_t FAR = KC + nbo;
        bool stop_zero = false;
        int npanels = 0;
        int64_t P = 0;
        while (nbo > 0 && kc < FAR && kr < m) {
            const int64_t cj = kc, rj = kr;
            int64_t w = FAR - cj;
            if (w > NBW) w = NBW;
            if (!ple && w > m - rj) w = m - rj;
            int32_t *mdst = bmap_dst + npanels * 64, *msrc = bmap_src + npanels * 64;
            PanelCtl *pctl = bctl + npanels;
            switch (RPT) {
            case 1: MODP_TRY((launch_panel<1, 32>(ctx, W, ldw, (int)m, (int)rj, (int)cj, (int)w, mp, ple, ipiv, pivval, pivinv, pivcol, colmap, mdst, msrc, pctl))); break;
            case 2: MODP_TRY((launch_panel<2, 32>(ctx, W, ldw, (int)m, (int)rj, (int)cj, (int)w, mp, ple, ipiv, pivval, pivinv, pivcol, colmap, mdst, msrc, pctl))); break;
            case 4: MODP_TRY((launch_panel<4, 16>(ctx, W, ldw, (int)m, (int)rj, (int)cj, (int)w, mp, ple, ipiv, pivval, pivinv, pivcol, colmap, mdst, msrc, pctl))); break;
            case 8: MODP_TRY((launch_panel<8, 8>(ctx, W, ldw, (int)m, (int)rj, (int)cj, (int)w, mp, ple, ipiv, pivval, pivinv, pivcol, colmap, mdst, msrc, pctl))); break;
            default: MODP_TRY((launch_panel<16, 4>(ctx, W, ldw, (int)m, (int)rj, (int)cj, (int)w, mp, ple, ipiv, pivval, pivinv, pivcol, colmap, mdst, msrc, pctl))); break;
            }
            ++npanels;
            HIP_TRY(ctx, hipMemcpyAsync(hctl, pctl, sizeof(PanelCtl), hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            const PanelCtl pc = *hctl;
            if (pc.status == 3)
                return set_error(ctx, MODP_ERR_NOT_INVERTIBLE, "pivot has no inverse modulo %lld (modulus not prime)",
                                 (long long)p);
            const int64_t npiv = pc.npiv;
            if (npiv > 0) {
                // the rest of the outer block: swaps; solve right of the panel
                const int64_t nblk = (FAR - KC) - w;
                if (nblk > 0) {
                    TimedLaunch tl(ctx, "elim_finish");
                    k_finish<<<grid_for(nblk * 64, 4096), 256, 0, ctx->stream>>>(
                        W, ldw, KC, FAR, (int)rj, (int)cj, (int)w, (int)KC, pctl, pivinv, pivcol, colmap, mdst, msrc,
                        Ubig, NBO, mp);
                    HIP_TRY(ctx, hipGetLastError());
                }
                // G1: rest of the outer block, every row below the pivots
                const int64_t r1 = rj + npiv;
                MODP_TRY(gemm_sub(r1, m - r1, cj + w, FAR - (cj + w), cj, w, Ubig + (cj - KC) + (cj + w - KC) * NBO,
                                  MODP_FORM_PLAIN_F64));
            }
            kr += npiv;
            kc += ple ? pc.ncols : npiv;
            P += npiv;
            if (pc.status == 1) { stop_zero = true; break; }
            if (pc.status == 2) break;
        }
        const int64_t consumed = kc - KC;
        if (P > 0) {
            // columns outside the block: the block's swaps; far columns: solve
            // with the block's L11, then the trailing GEMM with K = consumed
            const int64_t nout = n - nbo;
            if (nout > 0) {
                TimedLaunch tl(ctx, "elim_block_finish");
                int64_t g = (nout + 15) / 16;
                if (g > 2048) g = 2048;
                k_block_finish<<<(unsigned)g, 1024, 0, ctx->stream>>>(W, ldw, n, (int)KR, (int)KC, (int)FAR, (int)P,
                                                                      (int)consumed, npanels, bmap_dst, bmap_src,
                                                                      bctl, pivinv, pivcol, Ubig, NBO, mp);
                HIP_TRY(ctx, hipGetLastError());
            }
            const int64_t rbeg = KR + P;
            MODP_TRY(gemm_sub(rbeg, m - rbeg, FAR, n - FAR, KC, consumed, Ubig + (FAR - KC) * NBO, big_form));
        }
        if (!ple && kr < m && kc < n && (stop_zero || nbo == 0 || kc > last)) {
            // pluq zero column at kc (rows >= kr zero): whole block zero -> tail,
            // else move it to `last` and retry the same column
            HIP_TRY(ctx, hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
            k_any_nonzero<<<grid_for((m - kr) * (n - kc), 4096), 256, 0, ctx->stream>>>(W, ldw, kr, m, kc, n, flag);
            HIP_TRY(ctx, hipGetLastError());
            int *hflag = (int *)hctl;
            HIP_TRY(ctx, hipMemcpyAsync(hflag, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            if (!*hflag) {
                std::vector<int64_t> tail;
                for (int64_t c = kc; c < n; ++c) {
                    tail.push_back(c + 1);
                    tail.push_back(n);
                    col_stack.push_back(c + 1);
                    col_stack.push_back(n);
                }
                MODP_TRY(util_perm(ctx, EK_I32, 0, m, n, W, ldw, tail.data(), (int64_t)tail.size() / 2, 0));
                done = true;
            } else {
                int64_t sw[2] = {kc + 1, last + 1};
                col_stack.push_back(kc + 1);
                col_stack.push_back(last + 1);
                MODP_TRY(util_perm(ctx, EK_I32, 0, m, n, W, ldw, sw, 1, 0));
                --last;
            }
        }
        if (ple && kr >= m) break;
    }
    const int64_t rank = kr;

    // outputs
    if (U) {
        unsigned g = grid_for(m * n);
        if (kind == EK_F64) k_extract_u<double><<<g, 256, 0, ctx->stream>>>(m, n, W, ldw, (int)rank, pivcol, (double *)U, ldu);
        else k_extract_u<float><<<g, 256, 0, ctx->stream>>>(m, n, W, ldw, (int)rank, pivcol, (float *)U, ldu);
        HIP_TRY(ctx, hipGetLastError());
    }
    if (L && mode != ELIM_RREF) {
        unsigned g = grid_for(m * m);
        if (kind == EK_F64) k_extract_l<double><<<g, 256, 0, ctx->stream>>>(m, W, ldw, (int)rank, pivcol, pivval, (double *)L, ldl);
        else k_extract_l<float><<<g, 256, 0, ctx->stream>>>(m, W, ldw, (int)rank, pivcol, pivval, (float *)L, ldl);
        HIP_TRY(ctx, hipGetLastError());
    }
    std::vector<int32_t> hp((size_t)(2 * (rank > 0 ? rank : 1)));
    if (rank > 0) {
        HIP_TRY(ctx, hipMemcpyAsync(hp.data(), ipiv, sizeof(int32_t) * rank, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(hp.data() + rank, pivcol, sizeof(int32_t) * rank, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    res.rank = rank;
    for (int64_t i = 0; i < rank; ++i) {
        if (hp[i] != i) {
            res.row_swaps.push_back(i + 1);
            res.row_swaps.push_back((int64_t)hp[i] + 1);
        }
        res.pivcols.push_back((int64_t)hp[rank + i] + 1);
    }
    res.col_swaps = col_stack;
    if (mode == ELIM_LU) {
        res.perm.resize((size_t)m);
        for (int64_t i = 0; i < m; ++i) res.perm[i] = i + 1;
        for (int64_t i = 0; i < rank; ++i) std::swap(res.perm[i], res.perm[hp[i]]);
    }
    return MODP_OK;
}

// Back-elimination of a PLE-reduced working array into RREF: row blocks of
// 64 pivots from the bottom; each block is first reduced in place (unit upper
// triangular solve) and then eliminated from every row above it by a GEMM.
int elim_rref(modp_ctx *ctx, int form, int kind, int64_t m, int64_t n, const void *A, int64_t lda, void *R,
              int64_t ldr, int64_t p, ElimResult &res)
{
    // factor (PLE) into W without extracting L
    MODP_TRY(elim_run(ctx, ELIM_RREF, form, kind, m, n, A, lda, nullptr, 0, nullptr, 0, p, res));
    const ModP mp = make_modp(p);
    const int64_t rank = res.rank;
    int st = MODP_OK;
    const int64_t ldw = rup(m, 64);
    int32_t *W = (int32_t *)ws_get(ctx, WS_WORK, sizeof(int32_t) * (size_t)ldw * (size_t)n, &st);
    if (!W) return st;
    const int64_t mn = m < n ? m : n;
    const size_t npv = (size_t)mn + 64;
    int32_t *piv = (int32_t *)ws_get(ctx, WS_PIV, sizeof(int32_t) * (4 * npv + 4 * 64) + sizeof(PanelCtl) + 64, &st);
    if (!piv) return st;
    int32_t *pivcol = piv + 3 * npv;
    if (rank > 0) {
        k_zero_lower<<<grid_for(m * rank), 256, 0, ctx->stream>>>(W, ldw, m, (int)rank, pivcol);
        HIP_TRY(ctx, hipGetLastError());
    }
    if (rank > 1) {
        int32_t *G = (int32_t *)ws_get(ctx, WS_SCRATCH, sizeof(int32_t) * (size_t)(rank + 64) * 64, &st);
        if (!G) return st;
        const int bform = form == MODP_FORM_AUTO ? MODP_FORM_PLAIN_F64 : form;
        for (int64_t i1 = rank; i1 > 0;) {
            const int64_t i0 = i1 > 64 ? i1 - 64 : 0;
            const int nb = (int)(i1 - i0);
            std::vector<int32_t> pc0(1);
            HIP_TRY(ctx, hipMemcpyAsync(pc0.data(), pivcol + i0, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            const int64_t cs = pc0[0];
            // T = W[i0:i1, pivcol[i0:i1]] -> G (nb x nb), then reduce the block
            k_gather_cols<<<grid_for((int64_t)nb * nb), 256, 0, ctx->stream>>>(W, ldw, i0, nb, pivcol, (int)i0, nb, G, 64);
            HIP_TRY(ctx, hipGetLastError());
            k_trsm_upper_unit<<<grid_for((n - cs) * 64, 4096), 256, 0, ctx->stream>>>(W, ldw, n, (int)i0, nb, cs, G, 64, mp);
            HIP_TRY(ctx, hipGetLastError());
            if (i0 > 0) {
                // M = W[0:i0, pivcol[i0:i1]] -> G (i0 x nb); W[0:i0, cs:] -= M * W[i0:i1, cs:]
                k_gather_cols<<<grid_for(i0 * nb), 256, 0, ctx->stream>>>(W, ldw, 0, i0, pivcol, (int)i0, nb, G, i0);
                HIP_TRY(ctx, hipGetLastError());
                GemmArgs g;
                g.m = i0; g.n = n - cs; g.k = nb;
                g.A = G; g.lda = i0; g.ka = EK_I32;
                g.B = W + i0 + cs * ldw; g.ldb = ldw; g.kb = EK_I32;
                g.C = W + cs * ldw; g.ldc = ldw; g.kc = EK_I32;
                g.epi = EPI_SUB;
                g.mp = mp;
                if (bform == MODP_FORM_LIMB_I8 && nb >= 64) MODP_TRY(gemm_limb_i8(ctx, g));
                else MODP_TRY(gemm_plain_f64(ctx, g));
            }
            i1 = i0;
        }
    }
    unsigned g = grid_for(m * n);
    if (kind == EK_F64) k_extract_u<double><<<g, 256, 0, ctx->stream>>>(m, n, W, ldw, (int)rank, pivcol, (double *)R, ldr);
    else k_extract_u<float><<<g, 256, 0, ctx->stream>>>(m, n, W, ldw, (int)rank, pivcol, (float *)R, ldr);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MODP_OK;
}
