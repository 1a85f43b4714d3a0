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
//     k_panel    one 512-thread workgroup holds the panel in registers; per
//                column: wave64 argmax by DPP with lowest-row tie-break,
//                cross-wave through LDS, the pivot's inverse from a table (or
//                extended Euclid), and the rank-1 update of the remaining
//                panel columns; rows move only at the end, by a row map, and
//                the old values of the moved rows in the block's other
//                columns are staged.  Above 32 x 512 rows k_panel_g does the
//                same with the panel left in W (any row count).
//     k_update   the block's other columns: the row map, and right of the
//                panel the triangular solve with L11 (rows of U) and the
//                rank-w update of every row below the pivots.
//   after the outer block: k_block_finish applies the block's row maps and
//   the solve with the block's L11 to the far columns, then one GEMM with
//   K = NBO (limb or plain form) updates the far trailing block.
#include <stdlib.h>

#include <type_traits>

#include "common.h"

namespace {

// compile-time loop: f(integral_constant<int, J>) for J = B .. E-1
template <int B, int E> struct Unroll {
    template <class F> __device__ __forceinline__ static void run(F &f)
    {
        f(std::integral_constant<int, B>{});
        Unroll<B + 1, E>::run(f);
    }
};
template <int E> struct Unroll<E, E> {
    template <class F> __device__ __forceinline__ static void run(F &) {}
};

constexpr int NBO = 128;  // outer block width (K of the trailing GEMM)

struct PanelCtl {
    int npiv;    // pivots found
    int status;  // 0 all columns consumed, 1 zero column (pluq stop),
                 // 2 rows exhausted, 3 pivot not invertible
    int ncols;   // panel columns consumed
    int nmap;    // entries of the row gather map
    int kr0;     // first row of the panel
};

constexpr int PT = 512;  // threads of the panel workgroup (8 waves, 256-VGPR budget)

typedef int v2i_e __attribute__((ext_vector_type(2)));
typedef int v4i_e __attribute__((ext_vector_type(4)));

// VW consecutive int32 from a VW-aligned address, one vector load
template <int VW> __device__ __forceinline__ void load_rows(const int32_t *p, int32_t *x)
{
    if constexpr (VW == 4) {
        const v4i_e t = *reinterpret_cast<const v4i_e *>(p);
        x[0] = t[0]; x[1] = t[1]; x[2] = t[2]; x[3] = t[3];
    } else if constexpr (VW == 2) {
        const v2i_e t = *reinterpret_cast<const v2i_e *>(p);
        x[0] = t[0]; x[1] = t[1];
    } else {
        x[0] = p[0];
    }
}

// VW consecutive rows of one column; x[e * stride] is row e
template <int VW> __device__ __forceinline__ void store_rows(int32_t *p, const int32_t *x, int stride)
{
    if constexpr (VW == 4) {
        *reinterpret_cast<v4i_e *>(p) = v4i_e{x[0], x[stride], x[2 * stride], x[3 * stride]};
    } else if constexpr (VW == 2) {
        *reinterpret_cast<v2i_e *>(p) = v2i_e{x[0], x[stride]};
    } else {
        p[0] = x[0];
    }
}

template <bool SMALL> __device__ __forceinline__ uint32_t mulmod_t(uint32_t a, uint32_t b, const ModP &mp)
{
    if (SMALL) {
        // p < 2^16: operands and quotient fit 24 bits, so the full-rate
        // 24-bit multiplier gives the exact 32-bit products
        const uint32_t x = __umul24(a, b);
        const uint32_t q = (uint32_t)((float)x * mp.pinvf);
        int32_t r = (int32_t)(x - __umul24(q, mp.p32));
        r += r < 0 ? (int32_t)mp.p32 : 0;
        r -= r >= (int32_t)mp.p32 ? (int32_t)mp.p32 : 0;
        return (uint32_t)r;
    }
    return (uint32_t)dmod((double)a * (double)b, mp);
}

// x mod p for exact integers |x| <= 2^51: fl(x * pinv) is within
// |x| 2^-52 / p <= 1/(2p) of x/p, so q = rint(.) is within 1/2 + 1/(2p) < 1 of
// x/p, r = x - q p is exact with |r| < p, and one fix-up brings it to [0,p).
__device__ __forceinline__ double dmod_near(double x, const ModP &mp)
{
    const double r = fma(-rint(x * mp.pinv), mp.pd, x);
    return r < 0.0 ? r + mp.pd : r;
}

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Pivot-search keys: larger key = larger residue, then lower position.
template <bool SMALL> struct PivKey;
template <> struct PivKey<true> {
    typedef uint32_t T;  // residue < 2^16 in bits 30..15, 0x7fff - position below
    static __device__ __forceinline__ T make(uint32_t v, int pos) { return (v << 15) | (uint32_t)(0x7fff - pos); }
    static __device__ __forceinline__ uint32_t val(T k) { return k >> 15; }
    static __device__ __forceinline__ int pos(T k) { return 0x7fff - (int)(k & 0x7fffu); }
};
template <> struct PivKey<false> {
    typedef uint64_t T;
    static __device__ __forceinline__ T make(uint32_t v, int pos) { return ((uint64_t)v << 32) | (uint32_t)(0x7fffffff - pos); }
    static __device__ __forceinline__ uint32_t val(T k) { return (uint32_t)(k >> 32); }
    static __device__ __forceinline__ int pos(T k) { return 0x7fffffff - (int)(uint32_t)k; }
};

// Panel factorisation: one 512-thread workgroup, the panel W[kr0:m, c0:c0+w]
// in registers (thread t owns local rows t + 512*q) as exact integers in
// FP64.  Each rank-1 update is one FMA: for p < 2^16 every product is below
// 2^32, so up to 2^21 updates stay exact below 2^53 and only the pivot
// column and the pivot row are reduced (lazy reduction); larger p reduce
// after every update.
//
// One barrier per column: every wave finds its best candidate (largest
// residue, then lowest position) with DPP, and the lane holding it publishes
// that row, its index and its inverse in LDS before the barrier (double
// buffered by column parity); after the barrier all threads pick the winner
// among the waves and already have its row.  Rows never move: each row
// carries its current position (the reference's swap of positions pc and
// Ip), a row chosen as pivot is frozen with its normalised U row kept in
// LDS, and at the end every row is written to its final position.
template <int RPT, int NBW, bool SMALL>
__global__ __launch_bounds__(PT) void k_panel(int32_t *__restrict__ W, int64_t ldw, int m, int kr0_in, int c0,
                                              int w, ModP mp, int ple, const uint32_t *__restrict__ invtab,
                                              int32_t *__restrict__ ipiv, int32_t *__restrict__ pivval,
                                              int32_t *__restrict__ pivinv, int32_t *__restrict__ pivcol,
                                              int32_t *__restrict__ colmap, int32_t *__restrict__ map_dst,
                                              int32_t *__restrict__ map_src, PanelCtl *__restrict__ ctl,
                                              int *__restrict__ abort, int KC, int FAR, int32_t *__restrict__ stage,
                                              int *__restrict__ cursor)
{
    constexpr int NW = PT / 64;
    if (abort && *abort) {
        // an earlier panel of this block stopped: nothing to do
        if (threadIdx.x == 0) {
            ctl->npiv = 0;
            ctl->status = 4;
            ctl->ncols = 0;
            ctl->nmap = 0;
            ctl->kr0 = kr0_in;
        }
        return;
    }
    // the first row comes from the device cursor when the previous panel's
    // pivot count is not known to the host (lu, rref)
    const int kr0 = cursor ? *cursor : kr0_in;
    typedef PivKey<SMALL> K;
    typedef typename K::T key_t;
    __shared__ key_t s_key[2][NW];
    __shared__ double s_crow[2][NW][NBW];        // candidate rows (unreduced)
    __shared__ uint32_t s_cinv[2][NW];
    __shared__ int s_cidx[2][NW];                // candidate local row index
    __shared__ uint32_t s_U[NBW][NBW];
    __shared__ int s_pcol[NBW];
    // per pivot: source position, value, inverse; per column: pivot or -1
    // (kept in LDS during the steps, written out at the end)
    __shared__ int s_ip[NBW], s_pv[NBW], s_pi[NBW], s_cmap[NBW];
    __shared__ int s_cnt;
    __shared__ int s_msrc[2 * NBW];  // source row of each row-map entry

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rows = m - kr0;
    // Rows are read as VW-row vectors from the VW-aligned row kra <= kr0:
    // register row q of this thread is panel row rowi(q) (negative for the
    // sh rows above kr0, which only pad the first vector).  kra + RPT * PT >= m
    // because the host sizes RPT from m.
    constexpr int VW = RPT < 4 ? RPT : 4;
    const int kra = kr0 & ~(VW - 1), sh = kr0 - kra;
    auto rowi = [&](int q) { return VW * tid + (q % VW) + (q / VW) * (VW * PT) - sh; };
    double v[RPT][NBW];
    int pos[RPT];
    uint32_t pivmask = 0;
#pragma unroll
    for (int g = 0; g < RPT / VW; ++g) {
        const int64_t base = (int64_t)kra + VW * tid + g * (VW * PT);  // first row of the vector
#pragma unroll
        for (int j = 0; j < NBW; ++j) {
            int32_t x[VW];
#pragma unroll
            for (int e = 0; e < VW; ++e) x[e] = 0;
            if (j < w && base < ldw) load_rows<VW>(W + base + (int64_t)(c0 + j) * ldw, x);
#pragma unroll
            for (int e = 0; e < VW; ++e) {
                const int i = rowi(g * VW + e);
                // rows above kr0 keep their (final) values: the in-place
                // stores at the end write them back unchanged
                v[g * VW + e][j] = i < rows ? (double)x[e] : 0.0;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) pos[q] = rowi(q);
    if (tid == 0) s_cnt = 0;

    // rows that exist (bit q: register row q is a panel row)
    uint32_t live = 0;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int i = rowi(q);
        if (i >= 0 && i < rows) live |= 1u << q;
    }
    int pc = 0, status = 0, consumed = w, par = 0;
    bool active = true;
    // One step per panel column, unrolled at compile time so that v[][jc]
    // is a register.  Every step runs, also after the panel has stopped or
    // past its width, with a zero update: the step is straight-line code,
    // so the register array is never copied between branches.  Per-thread
    // work is kept to the rows the thread owns: the pivot row is published
    // unreduced, and its reduction and normalisation are spread over the
    // lanes (lane j handles column j) and broadcast back with readlane.
    auto step = [&](auto JC) __attribute__((always_inline)) {
        constexpr int jc = decltype(JC)::value;
        const bool run = active && jc < w && pc < rows;
        if (active && jc < w && pc >= rows) {
            status = 2;
            consumed = jc;
            active = false;
        }
        // this thread's best row among those not yet pivots; column jc is
        // final after this step, so it is reduced here once and for all
        const uint32_t elig = run ? live & ~pivmask : 0u;
        key_t mine = 0;
        int mq = 0;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            if (SMALL) v[q][jc] = dmod_near(v[q][jc], mp);
            // selects, not branches: keys of rows that are not eligible are 0
            const key_t k = ((elig >> q) & 1u) ? K::make((uint32_t)v[q][jc], pos[q]) : (key_t)0;
            const bool better = k > mine;
            mine = better ? k : mine;
            mq = better ? q : mq;
        }
        // every lane fetches the inverse of its own candidate before the wave
        // reduction, so the table load overlaps it
        const uint32_t myinv = (invtab && K::val(mine) != 0) ? invtab[K::val(mine)] : 0u;
        const key_t wbest = wave_max(mine);
        // the lane holding the wave's candidate (keys are unique) publishes
        // its index and inverse; its register row is made wave-uniform, so
        // the row is published through a scalar branch per register row
        // rather than a per-lane select over all of them
        const uint64_t wmask = __ballot(mine == wbest && K::val(wbest) != 0);
        if (wmask) {
            const int wl = (int)__builtin_ctzll(wmask);
            const int wq = __builtin_amdgcn_readlane(mq, wl);
            if (lane == wl) {
                s_cinv[par][wave] = invtab ? myinv : modinv_fast(K::val(wbest), mp);
                s_cidx[par][wave] = rowi(mq);
            }
            // one divergent branch (the winning lane), then scalar branches
            // on the wave-uniform register row
            auto publish = [&](auto Q) __attribute__((always_inline)) {
                constexpr int q = decltype(Q)::value;
                if (wq == q) {
#pragma unroll
                    for (int j = 0; j < NBW; ++j)
                        if (j > jc) s_crow[par][wave][j] = v[q][j];
                    // keeps the cases' stores apart: merged, they would
                    // become one store of an indexed (scratch) register row
                    asm volatile("" ::: "memory");
                }
            };
            if (lane == wl) Unroll<0, RPT>::run(publish);
        }
        if (lane == 0) s_key[par][wave] = wbest;
        __syncthreads();
        key_t best = s_key[par][0];
        int bw = 0;
#pragma unroll
        for (int x = 1; x < NW; ++x) {
            const key_t k = s_key[par][x];
            if (k > best) { best = k; bw = x; }
        }
        bw = uniform(bw);
        const int cp = par;
        par ^= 1;
        const uint32_t V = (uint32_t)uniform((int)K::val(best));
        const int Ip = uniform(K::pos(best));
        const uint32_t inv = V != 0 ? (uint32_t)uniform((int)s_cinv[cp][bw]) : 0u;
        const int wrow = V != 0 ? uniform(s_cidx[cp][bw]) : -1;
        const bool piv = run && V != 0 && inv != 0;
        if (run && V == 0) {
            if (ple) {
                if (tid == 0) s_cmap[jc] = -1;
            } else {
                status = 1;
                consumed = jc;
                active = false;
            }
        }
        if (run && V != 0 && inv == 0) {
            status = 3;
            consumed = jc;
            active = false;
        }
        // positions: the winner moves to pc, the row at pc moves to Ip
        if (piv) {
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const bool win = rowi(q) == wrow;
                pivmask |= win ? 1u << q : 0u;
                pos[q] = win ? pc : (pos[q] == pc ? Ip : pos[q]);
            }
        }
        // normalised pivot row: lane j reduces and scales column j, wave 0
        // keeps it for the output, readlane hands every lane the values
        uint32_t ul = 0;
        if (piv && lane > jc && lane < NBW) {
            const double x = s_crow[cp][bw][lane];
            const uint32_t r = (uint32_t)(SMALL ? dmod_near(x, mp) : x);
            ul = mulmod_t<SMALL>(r, inv, mp);
            if (wave == 0) s_U[pc][lane] = ul;
        }
        double u[NBW];
#pragma unroll
        for (int j = 0; j < NBW; ++j) u[j] = j > jc ? (double)(uint32_t)__builtin_amdgcn_readlane((int)ul, j) : 0.0;
        // rank-1 update of the rows that are not pivots (a zero update when
        // this step found no pivot)
        const uint32_t upd = live & ~pivmask;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const double l = ((upd >> q) & 1u) ? v[q][jc] : 0.0;
#pragma unroll
            for (int j = 0; j < NBW; ++j)
                if (j > jc) {
                    v[q][j] = fma(-l, u[j], v[q][j]);
                    if (!SMALL) v[q][j] = dmod(v[q][j], mp);
                }
        }
        if (piv && tid == 0) {
            s_ip[pc] = Ip;
            s_pv[pc] = (int)V;
            s_pi[pc] = (int)inv;
            s_pcol[pc] = jc;
            s_cmap[jc] = pc;
        }
        pc += piv ? 1 : 0;
    };
    Unroll<0, NBW>::run(step);
    __syncthreads();

    // Row gather map: positions 0..pc-1 first, in order, then every other
    // row that moved.
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int i = rowi(q);
        if (i >= 0 && i < rows) {
            const int dst = kr0 + pos[q];
            if (pos[q] < pc) {
                map_dst[pos[q]] = dst;
                map_src[pos[q]] = kr0 + i;
                s_msrc[pos[q]] = kr0 + i;
            } else if (pos[q] != i) {
                const int e = pc + atomicAdd(&s_cnt, 1);
                map_dst[e] = dst;
                map_src[e] = kr0 + i;
                s_msrc[e] = kr0 + i;
            }
        }
    }
    __syncthreads();
    // The old values of every moved row in the block's other columns go to
    // the stage (stage[e + 64 ci], ci counting the block columns left of the
    // panel, then those right of it) for the update kernel; the loads are
    // issued here and stored after the panel's own stores.
    constexpr int SPT = (2 * NBW * NBO + PT - 1) / PT;
    const int nmap_all = pc + s_cnt, ncol_other = (FAR - KC) - w, nstage = nmap_all * ncol_other;
    int32_t sv[SPT];
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
        const int idx = tid + k * PT;
        if (idx < nstage) {
            const int e = idx % nmap_all, ci = idx / nmap_all;
            const int c = ci < c0 - KC ? KC + ci : KC + ci + w;
            sv[k] = W[(int64_t)s_msrc[e] + (int64_t)c * ldw];
        }
    }

    // Every row to its final position, in two passes: first every row is
    // stored in place with one vector store per VW rows and column (no
    // per-row branches), then, once those stores are complete, the rows
    // that moved and the pivot rows (1 at the pivot column, the normalised
    // U row after it) overwrite their destinations.  The rows that moved
    // leave exactly the positions that other moved rows fill, so every
    // in-place value that is wrong is overwritten.
    {
        // the values are converted first and the stores issued back to back
        // (full-width panels need no per-column test)
        const int64_t basev = (int64_t)kra + VW * tid;
        int32_t out[RPT][NBW];
#pragma unroll
        for (int q = 0; q < RPT; ++q)
#pragma unroll
            for (int j = 0; j < NBW; ++j) out[q][j] = (int32_t)(uint32_t)v[q][j];
#pragma unroll
        for (int g = 0; g < RPT / VW; ++g) {
            const int64_t base = basev + g * (VW * PT);
            int32_t *pg = W + base + (int64_t)c0 * ldw;
            if (base + VW <= (int64_t)m) {
                if (w == NBW) {
#pragma unroll
                    for (int j = 0; j < NBW; ++j) store_rows<VW>(pg + (int64_t)j * ldw, &out[g * VW][j], NBW);
                } else {
#pragma unroll
                    for (int j = 0; j < NBW; ++j)
                        if (j < w) store_rows<VW>(pg + (int64_t)j * ldw, &out[g * VW][j], NBW);
                }
            }
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int i = rowi(q);
        const bool pv = (pivmask >> q) & 1u;
        const bool tail = i >= 0 && i < rows && kra + VW * tid + (q / VW) * (VW * PT) + VW > m;
        if (i >= 0 && i < rows && (pv || pos[q] != i || tail)) {
            const int64_t dst = (int64_t)kr0 + pos[q];
            const int k = pos[q], jk = pv ? s_pcol[k] : NBW;
#pragma unroll
            for (int j = 0; j < NBW; ++j)
                if (j < w)
                    W[dst + (int64_t)(c0 + j) * ldw] =
                        (int32_t)(j < jk ? (uint32_t)v[q][j] : (j == jk ? 1u : s_U[k][j]));
        }
    }
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
        const int idx = tid + k * PT;
        if (idx < nstage) stage[idx % nmap_all + 64 * (idx / nmap_all)] = sv[k];
    }
    if (tid < pc) {
        ipiv[kr0 + tid] = kr0 + s_ip[tid];
        pivval[kr0 + tid] = s_pv[tid];
        pivinv[kr0 + tid] = s_pi[tid];
        pivcol[kr0 + tid] = c0 + s_pcol[tid];
    }
    if (tid < w) colmap[tid] = tid < consumed ? s_cmap[tid] : -1;
    if (tid == 0) {
        ctl->npiv = pc;
        ctl->status = status;
        ctl->ncols = consumed;
        ctl->nmap = pc + s_cnt;
        ctl->kr0 = kr0;
        if (cursor) *cursor = kr0 + pc;
        if (abort && status != 0) *abort = 1;
    }
}

// Panel factorisation for any number of rows (the register panel holds at
// most 32 x 512 rows and its small-prime key 2^15 positions): the same
// outputs as k_panel from one 1024-thread workgroup that keeps the panel in
// W.  Thread t owns local rows t + 1024 q and is the only thread that reads
// or writes them during the column steps; a row's current position (the
// reference's swap of positions pc and Ip) is kept in gpos[row], the row at
// each position in grow[pos].  Per column: the owners scan their eligible
// rows (position >= pc) with 64-bit keys (residue, then lowest position),
// the wave winners publish their rows and inverses in LDS, and after one
// barrier every thread knows the pivot; thread 0 swaps the positions, the
// lanes below w normalise the pivot row, and after a second barrier the
// owners apply the rank-1 update to their rows (residues reduced every
// time).  At the end the moved rows (at most 2 w) go through LDS to their
// positions, the pivot rows with their 1 and normalised U row.
constexpr int GPT = 1024, GNBW = 16;
template <int NBW>
__global__ __launch_bounds__(GPT) void k_panel_g(int32_t *__restrict__ W, int64_t ldw, int m, int kr0_in, int c0,
                                                 int w, ModP mp, int ple, const uint32_t *__restrict__ invtab,
                                                 int32_t *__restrict__ ipiv, int32_t *__restrict__ pivval,
                                                 int32_t *__restrict__ pivinv, int32_t *__restrict__ pivcol,
                                                 int32_t *__restrict__ colmap, int32_t *__restrict__ map_dst,
                                                 int32_t *__restrict__ map_src, PanelCtl *__restrict__ ctl,
                                                 int *__restrict__ abort, int KC, int FAR, int32_t *__restrict__ stage,
                                                 int *__restrict__ cursor, int32_t *__restrict__ gpos,
                                                 int32_t *__restrict__ grow)
{
    constexpr int NW = GPT / 64;
    if (abort && *abort) {
        if (threadIdx.x == 0) {
            ctl->npiv = 0;
            ctl->status = 4;
            ctl->ncols = 0;
            ctl->nmap = 0;
            ctl->kr0 = kr0_in;
        }
        return;
    }
    const int kr0 = cursor ? *cursor : kr0_in;
    __shared__ uint64_t s_key[NW];
    __shared__ int s_row[NW];
    __shared__ uint32_t s_cinv[NW];
    __shared__ uint32_t s_crow[NW][NBW];
    __shared__ uint32_t s_U[NBW][NBW];
    __shared__ int s_ip[NBW], s_pv[NBW], s_pi[NBW], s_pcol[NBW], s_cmap[NBW];
    __shared__ int s_msrc[2 * NBW], s_mdst[2 * NBW];
    __shared__ uint32_t s_mv[2 * NBW][NBW];
    __shared__ int s_cnt;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rows = m - kr0;
    int32_t *P = W + kr0 + (int64_t)c0 * ldw;  // local row i, panel column j: P[i + j ldw]
    for (int i = tid; i < rows; i += GPT) {
        gpos[i] = i;
        grow[i] = i;
    }
    if (tid == 0) s_cnt = 0;
    __syncthreads();

    int pc = 0, status = 0, consumed = w;
    bool active = true;
    for (int jc = 0; jc < NBW; ++jc) {
        const bool run = active && jc < w && pc < rows;
        if (active && jc < w && pc >= rows) {
            status = 2;
            consumed = jc;
            active = false;
        }
        uint64_t mine = 0;
        int mrow = -1;
        if (run) {
            for (int i = tid; i < rows; i += GPT) {
                const int ps = gpos[i];
                if (ps < pc) continue;
                const uint32_t v = (uint32_t)P[i + (int64_t)jc * ldw];
                const uint64_t k = ((uint64_t)v << 32) | (uint32_t)(0x7fffffff - ps);
                if (v != 0 && k > mine) {
                    mine = k;
                    mrow = i;
                }
            }
        }
        const uint64_t wbest = wave_max(mine);
        if (mine != 0 && mine == wbest) {
            // keys are unique: one lane per wave publishes its row
            const uint32_t V = (uint32_t)(mine >> 32);
            s_key[wave] = mine;
            s_row[wave] = mrow;
            s_cinv[wave] = invtab ? invtab[V] : modinv_fast(V, mp);
            for (int j = jc + 1; j < NBW; ++j) s_crow[wave][j] = j < w ? (uint32_t)P[mrow + (int64_t)j * ldw] : 0u;
        }
        if (lane == 0 && wbest == 0) s_key[wave] = 0;
        __syncthreads();
        uint64_t best = s_key[0];
        int bw = 0;
        for (int x = 1; x < NW; ++x) {
            const uint64_t k = s_key[x];
            if (k > best) {
                best = k;
                bw = x;
            }
        }
        const uint32_t V = (uint32_t)(best >> 32);
        const int Ip = 0x7fffffff - (int)(uint32_t)best;
        const uint32_t inv = V != 0 ? s_cinv[bw] : 0u;
        const int wrow = V != 0 ? s_row[bw] : -1;
        const bool piv = run && V != 0 && inv != 0;
        if (run && V == 0) {
            if (ple) {
                if (tid == 0) s_cmap[jc] = -1;
            } else {
                status = 1;
                consumed = jc;
                active = false;
            }
        }
        if (run && V != 0 && inv == 0) {
            status = 3;
            consumed = jc;
            active = false;
        }
        if (piv) {
            if (tid == 0) {
                // the winner moves to position pc, the row at pc to Ip
                const int rpc = grow[pc];
                gpos[wrow] = pc;
                gpos[rpc] = Ip;
                grow[Ip] = rpc;
                grow[pc] = wrow;
                s_ip[pc] = Ip;
                s_pv[pc] = (int)V;
                s_pi[pc] = (int)inv;
                s_pcol[pc] = jc;
                s_cmap[jc] = pc;
            }
            if (tid > jc && tid < NBW) s_U[pc][tid] = tid < w ? mulmod(s_crow[bw][tid], inv, mp) : 0u;
        }
        __syncthreads();
        if (piv) {
            const int pn = pc + 1;
            for (int i = tid; i < rows; i += GPT) {
                if (gpos[i] < pn) continue;
                const uint32_t l = (uint32_t)P[i + (int64_t)jc * ldw];
                if (l == 0) continue;
                for (int j = jc + 1; j < w; ++j) {
                    int32_t *x = P + i + (int64_t)j * ldw;
                    *x = (int32_t)submod((uint32_t)*x, mulmod(l, s_U[pc][j], mp), mp);
                }
            }
            pc = pn;
        }
        // the next step's LDS writes come after every read of this one
        __syncthreads();
    }

    // Row gather map: positions 0..pc-1 first, in order, then every other
    // row that moved.
    for (int i = tid; i < rows; i += GPT) {
        const int ps = gpos[i];
        if (ps < pc) {
            map_dst[ps] = kr0 + ps;
            map_src[ps] = kr0 + i;
            s_msrc[ps] = i;
            s_mdst[ps] = ps;
        } else if (ps != i) {
            const int e = pc + atomicAdd(&s_cnt, 1);
            map_dst[e] = kr0 + ps;
            map_src[e] = kr0 + i;
            s_msrc[e] = i;
            s_mdst[e] = ps;
        }
    }
    __syncthreads();
    const int nmap_all = pc + s_cnt;
    // old values of the moved rows in the block's other columns
    const int ncol_other = (FAR - KC) - w, nstage = nmap_all * ncol_other;
    for (int idx = tid; idx < nstage; idx += GPT) {
        const int e = idx % nmap_all, ci = idx / nmap_all;
        const int c = ci < c0 - KC ? KC + ci : KC + ci + w;
        stage[e + 64 * ci] = W[(int64_t)kr0 + s_msrc[e] + (int64_t)c * ldw];
    }
    // the moved rows' final panel values, then their stores
    for (int idx = tid; idx < nmap_all * NBW; idx += GPT) {
        const int e = idx % nmap_all, j = idx / nmap_all;
        if (j < w) {
            const int k = s_mdst[e];
            const int jk = k < pc ? s_pcol[k] : NBW;
            s_mv[e][j] = j < jk ? (uint32_t)P[s_msrc[e] + (int64_t)j * ldw] : (j == jk ? 1u : s_U[k][j]);
        }
    }
    __syncthreads();
    for (int idx = tid; idx < nmap_all * NBW; idx += GPT) {
        const int e = idx % nmap_all, j = idx / nmap_all;
        if (j < w) P[s_mdst[e] + (int64_t)j * ldw] = (int32_t)s_mv[e][j];
    }
    if (tid < pc) {
        ipiv[kr0 + tid] = kr0 + s_ip[tid];
        pivval[kr0 + tid] = s_pv[tid];
        pivinv[kr0 + tid] = s_pi[tid];
        pivcol[kr0 + tid] = c0 + s_pcol[tid];
    }
    if (tid < w) colmap[tid] = tid < consumed ? s_cmap[tid] : -1;
    if (tid == 0) {
        ctl->npiv = pc;
        ctl->status = status;
        ctl->ncols = consumed;
        ctl->nmap = nmap_all;
        ctl->kr0 = kr0;
        if (cursor) *cursor = kr0 + pc;
        if (abort && status != 0) *abort = 1;
    }
}

// After panel (kr0, c0, w) of the outer block [KC, FAR), in one launch:
// the panel's row moves for every other column of the block, and for the
// columns right of the panel the triangular solve with L11 (their rows of U)
// and the rank-npiv update of every row below the pivots.  The old values
// of moved rows come from the stage the panel kernel filled before anything
// moved, so no block reads a row that another block writes.  Grid: x =
// 64-row tiles from kr0 down, y = 8-column groups, the groups left of the
// panel first (moves only, done by the blocks of the first row tile).
// LAZY (p < 2^16): at most 32 products below 2^32 are summed exactly in FP64
// before one reduction; otherwise every product is reduced.
constexpr int UT_R = 64, UT_C = 8;
template <int WMAX, bool LAZY>
__global__ __launch_bounds__(256) void k_update(int32_t *__restrict__ W, int64_t ldw, int m, int KC, int FAR,
                                                int c0, int w, const PanelCtl *__restrict__ ctl,
                                                const int32_t *__restrict__ pivinv,
                                                const int32_t *__restrict__ pivcol,
                                                const int32_t *__restrict__ map_dst,
                                                const int32_t *__restrict__ stage, ModP mp,
                                                const int *__restrict__ skip)
{
    // the skip flag and the panel's counts are read together
    const int skipped = skip ? *skip : 0;
    const int npiv = ctl->npiv, nmap = ctl->nmap, kr0 = ctl->kr0;
    if (skipped || nmap == 0) return;
    constexpr int CPT = UT_C / 4;            // columns per thread
    __shared__ uint32_t Lf[WMAX][WMAX + 1];  // the panel's top w x w block
    __shared__ uint32_t Ls[WMAX][WMAX + 1];  // L11 (row i, pivot j < i)
    __shared__ uint32_t invs[WMAX];
    __shared__ int s_pj[WMAX];              // panel column of pivot i
    __shared__ int s_dst[64];               // row-map destinations
    __shared__ double Ut[UT_C][WMAX];        // U by column, then panel column (0: no pivot)
    __shared__ int s_e[UT_R];               // tile row -> row-map entry, or -1
    const int tid = threadIdx.x, cl = tid >> 6;
    const int nl = c0 - KC, nr = FAR - (c0 + w);
    const int gl = (nl + UT_C - 1) / UT_C;
    const int g = blockIdx.y;
    if (g < gl) {
        // left of the panel: the row moves only
        if (blockIdx.x != 0) return;
        const int cb = g * UT_C, nc = min(UT_C, nl - cb);
        for (int idx = tid; idx < nmap * nc; idx += 256) {
            const int e = idx % nmap, ci = cb + idx / nmap;
            W[(int64_t)map_dst[e] + (int64_t)(KC + ci) * ldw] = stage[e + 64 * ci];
        }
        return;
    }
    const int cb = (g - gl) * UT_C, nc = min(UT_C, nr - cb);
    const int ci0 = nl + cb;        // stage column of the group's first column
    const int cfirst = c0 + w + cb;  // its column in W
    const int64_t r0 = (int64_t)kr0 + (int64_t)blockIdx.x * UT_R;
    if (r0 >= m) return;  // the grid covers the most rows the panel can have
    const int64_t r = r0 + (tid & 63);
    const bool row_ok = r < m;

    // every global load is issued before anything waits for one: the
    // panel's top block, pivot columns and inverses, the row map, the
    // solve's inputs (thread k < nc: column k of the pivot rows' old
    // values), this thread's row of L and its old values.  Indices are
    // clamped into range so that the loads are unconditional; values that
    // are not needed are masked when they are used.
    constexpr int FPT = (WMAX * WMAX + 255) / 256;
    const int wl = w - 1;
    const int64_t rr = row_ok ? r : (int64_t)kr0;
    int32_t lf[FPT];
#pragma unroll
    for (int k = 0; k < FPT; ++k) {
        const int idx = tid + 256 * k, i = min(idx / WMAX, wl), jj = min(idx % WMAX, wl);
        lf[k] = W[(int64_t)(kr0 + i) + (int64_t)(c0 + jj) * ldw];
    }
    const int tw = min(tid, WMAX - 1);
    const int32_t iv = pivinv[kr0 + tw];
    const int32_t pjv = pivcol[kr0 + tw];
    const int32_t mdv = map_dst[min(tid, 63)];
    int32_t xs0[WMAX];
    const int64_t xcol = 64 * (int64_t)(ci0 + min(tid, nc - 1));
#pragma unroll
    for (int i = 0; i < WMAX; ++i) xs0[i] = stage[min(i, wl) + xcol];
    int32_t lr0[WMAX];
#pragma unroll
    for (int jj = 0; jj < WMAX; ++jj) lr0[jj] = W[rr + (int64_t)(c0 + min(jj, wl)) * ldw];
    int32_t cv[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) cv[k] = W[rr + (int64_t)(cfirst + min(cl + 4 * k, nc - 1)) * ldw];
    for (int idx = tid; idx < WMAX * UT_C; idx += 256) Ut[idx / WMAX][idx % WMAX] = 0.0;
    if (tid < UT_R) s_e[tid] = -1;
#pragma unroll
    for (int k = 0; k < FPT; ++k) {
        const int idx = tid + 256 * k;
        if (idx < WMAX * WMAX) Lf[idx / WMAX][idx % WMAX] = (idx / WMAX < w && idx % WMAX < w) ? (uint32_t)lf[k] : 0u;
    }
    if (tid < WMAX) {
        invs[tid] = (uint32_t)iv;
        s_pj[tid] = pjv - c0;
    }
    if (tid < 64) s_dst[tid] = mdv;
    uint32_t xs[WMAX];
#pragma unroll
    for (int i = 0; i < WMAX; ++i) xs[i] = i < w ? (uint32_t)xs0[i] : 0u;
    double Lr[WMAX];
#pragma unroll
    for (int jj = 0; jj < WMAX; ++jj) Lr[jj] = jj < w ? (double)lr0[jj] : 0.0;
    __syncthreads();
    if (tid >= npiv && tid < nmap) {
        const int64_t d = (int64_t)s_dst[tid] - r0;
        if (d >= 0 && d < UT_R) s_e[d] = tid;
    }
    for (int idx = tid; idx < WMAX * WMAX; idx += 256) {
        const int i = idx / WMAX, j = idx % WMAX;
        Ls[i][j] = (j < i && i < npiv) ? Lf[i][s_pj[j]] : 0u;
    }
    __syncthreads();
    // rows of U for the group's columns: thread k solves column k, right
    // looking (each solved value is subtracted from the rows below at once,
    // so the dependent chain per row is one normalisation and one FMA)
    if (tid < nc) {
        double a[WMAX];
#pragma unroll
        for (int i = 0; i < WMAX; ++i) a[i] = (double)xs[i];
#pragma unroll
        for (int j = 0; j < WMAX; ++j) {
            if (j < npiv) {
                // x_j = a_j / pivot_j mod p; for small p in FP64 throughout
                // (the reduced a_j times the inverse is below 2^32)
                const double x = LAZY ? dmod_near(dmod_near(a[j], mp) * (double)invs[j], mp)
                                      : (double)mulmod((uint32_t)a[j], invs[j], mp);
                Ut[tid][s_pj[j]] = x;
#pragma unroll
                for (int i = j + 1; i < WMAX; ++i) {
                    a[i] = fma(-(double)Ls[i][j], x, a[i]);
                    if (!LAZY) a[i] = dmod(a[i], mp);
                }
            }
        }
    }
    __syncthreads();
    if (!row_ok) return;
    const int rel = (int)(r - kr0);
    if (rel < npiv) {
        // pivot row: its row of U
        const int pj = s_pj[rel];
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int col = cl + 4 * k;
            if (col < nc) W[r + (int64_t)(cfirst + col) * ldw] = (int32_t)(uint32_t)Ut[col][pj];
        }
        return;
    }
    const int e = s_e[tid & 63];
    if (e >= 0) {
        // a row that moved here: its old values are in the stage
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int col = cl + 4 * k;
            if (col < nc) cv[k] = stage[e + 64 * (ci0 + col)];
        }
    }
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int col = cl + 4 * k;
        if (col < nc) {
            double acc = (double)cv[k];
#pragma unroll
            for (int jj = 0; jj < WMAX; ++jj) {
                acc = fma(-Lr[jj], Ut[col][jj], acc);
                if (!LAZY) acc = dmod(acc, mp);
            }
            W[r + (int64_t)(cfirst + col) * ldw] = (int32_t)(uint32_t)(LAZY ? dmod_near(acc, mp) : acc);
        }
    }
}

constexpr int MAXPAN = 64;  // inner panels per outer block (NBO / smallest panel width, 2)

// End of an outer block with P pivots (rows [KR, KR+P)), whose inner panels
// left their row gather maps in bmap_*: one wave per column outside the block
// [KC, FAR).  Columns left of the block get the swaps; far columns get the
// swaps, then the solve with the block's L11 (P x P, in LDS; lanes hold rows
// i and i+64), and their U rows go to Ubig (rows 0..consumed-1 indexed by
// block column, zero for block columns without a pivot).
template <bool LAZY>
__global__ __launch_bounds__(1024) void k_block_finish(int32_t *__restrict__ W, int64_t ldw, int64_t n, int KR,
                                                       int KC, int FAR, int P, int consumed, int npanels,
                                                       const int32_t *__restrict__ bmap_dst,
                                                       const int32_t *__restrict__ bmap_src,
                                                       const PanelCtl *__restrict__ bctl,
                                                       const int32_t *__restrict__ pivinv,
                                                       const int32_t *__restrict__ pivcol,
                                                       int32_t *__restrict__ Ubig, int64_t ldub, ModP mp,
                                                       const int *__restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ uint32_t Ls[128][129];
    __shared__ uint32_t invs[128];
    __shared__ int ucolmap[128];
    __shared__ int s_pc[128];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < 128) {
        const int t = min(tid, P - 1);
        s_pc[tid] = pivcol[KR + t];
        invs[tid] = (uint32_t)pivinv[KR + t];
        ucolmap[tid] = -1;
    }
    __syncthreads();
    // L11 into LDS: lanes take consecutive rows of one pivot column, and
    // every load is issued before the first LDS store
    {
        constexpr int LPT = 128 * 128 / 1024;
        uint32_t lv[LPT];
#pragma unroll
        for (int k = 0; k < LPT; ++k) {
            const int idx = tid + 1024 * k, i = min(idx % 128, P - 1), j = min(idx / 128, P - 1);
            lv[k] = (uint32_t)W[(int64_t)(KR + i) + (int64_t)s_pc[j] * ldw];
        }
        // rows scaled by their pivot's inverse: the system becomes unit
        // lower triangular, so the substitution needs no multiply per step
#pragma unroll
        for (int k = 0; k < LPT; ++k) {
            const int idx = tid + 1024 * k, i = idx % 128, j = idx / 128;
            Ls[i][j] = (j < i && i < P) ? mulmod(lv[k], invs[i], mp) : 0u;
        }
    }
    if (tid < P) ucolmap[s_pc[tid] - KC] = tid;
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
        // forward substitution with the unit triangular system (rows and
        // right-hand side scaled by the pivots' inverses), lanes hold rows
        // lane and lane + 64: at step j every lane reduces its own sum (only
        // lane j's is used), readlane broadcasts it, and the rows below
        // subtract L[i][j] x_j.  LAZY (p < 2^16): the sums of at most 127
        // products below 2^32 stay exact in FP64 and are reduced once.
        double a0 = lane < P ? (double)mulmod((uint32_t)col[KR + lane], invs[lane], mp) : 0.0;
        double a1 = lane + 64 < P ? (double)mulmod((uint32_t)col[KR + lane + 64], invs[lane + 64], mp) : 0.0;
        for (int j = 0; j < P; ++j) {
            const double aj = j < 64 ? a0 : a1;
            const uint32_t xl = (uint32_t)(LAZY ? dmod_near(aj, mp) : aj);
            const uint32_t xj = (uint32_t)__builtin_amdgcn_readlane((int)xl, j & 63);
            const double xd = (double)xj;
            if (j < 64) {
                a0 = lane == j ? xd : a0;
                if (lane > j) {
                    a0 = fma(-(double)Ls[lane][j], xd, a0);
                    if (!LAZY) a0 = dmod(a0, mp);
                }
            } else if (lane == j - 64) {
                a1 = xd;
            }
            if (lane + 64 > j) {
                a1 = fma(-(double)Ls[lane + 64][j], xd, a1);
                if (!LAZY) a1 = dmod(a1, mp);
            }
        }
        const uint32_t x0 = (uint32_t)a0, x1 = (uint32_t)a1;
        if (lane < P) col[KR + lane] = (int32_t)x0;
        if (lane + 64 < P) col[KR + lane + 64] = (int32_t)x1;
        // Ubig rows 0..consumed-1 of this column: pivot i's value or zero
        int32_t *ucol = Ubig + (c - KC) * ldub;
        const int a0i = um0 < 0 ? 0 : um0, a1i = um1 < 0 ? 0 : um1;
        const uint32_t y00 = (uint32_t)__shfl((int)x0, a0i & 63), y01 = (uint32_t)__shfl((int)x1, a0i & 63);
        const uint32_t y10 = (uint32_t)__shfl((int)x0, a1i & 63), y11 = (uint32_t)__shfl((int)x1, a1i & 63);
        if (lane < consumed) ucol[lane] = um0 < 0 ? 0 : (int32_t)(a0i < 64 ? y00 : y01);
        if (lane + 64 < consumed) ucol[lane + 64] = um1 < 0 ? 0 : (int32_t)(a1i < 64 ? y10 : y11);
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
    {
        // every load issued before the first LDS store
        constexpr int TPT = 64 * 64 / 256;
        int32_t tv[TPT];
#pragma unroll
        for (int k = 0; k < TPT; ++k) {
            const int idx = tid + 256 * k, i = min(idx % 64, nb - 1), j = min(idx / 64, nb - 1);
            tv[k] = G[i + j * ldg];
        }
#pragma unroll
        for (int k = 0; k < TPT; ++k) {
            const int idx = tid + 256 * k, i = idx % 64, j = idx / 64;
            Ts[i][j] = (i < nb && j < nb) ? (uint32_t)tv[k] : 0u;
        }
    }
    __syncthreads();
    for (int64_t c = cs + (int64_t)blockIdx.x * 4 + wave; c < n; c += (int64_t)gridDim.x * 4) {
        int32_t *col = W + c * ldw + i0;
        uint32_t x = lane < nb ? (uint32_t)col[lane] : 0u;
        for (int j = nb - 1; j > 0; --j) {
            const uint32_t xj = (uint32_t)__builtin_amdgcn_readlane((int)x, j);
            if (lane < j) x = submod(x, mulmod(Ts[lane][j], xj, mp), mp);
        }
        if (lane < nb) col[lane] = (int32_t)x;
    }
}

template <int RPT, int NBW>
int launch_panel(modp_ctx *ctx, int32_t *W, int64_t ldw, int m, int kr0, int c0, int w, const ModP &mp, int ple,
                 const uint32_t *invtab, int32_t *ipiv, int32_t *pivval, int32_t *pivinv, int32_t *pivcol,
                 int32_t *colmap, int32_t *map_dst, int32_t *map_src, PanelCtl *ctl, int *abort, int KC, int FAR,
                 int32_t *stage, int *cursor)
{
    TimedLaunch tl(ctx, "elim_panel");
    const size_t shm = 0;
    if (mp.small)
        k_panel<RPT, NBW, true><<<1, PT, shm, ctx->stream>>>(W, ldw, m, kr0, c0, w, mp, ple, invtab, ipiv, pivval,
                                                            pivinv, pivcol, colmap, map_dst, map_src, ctl, abort,
                                                            KC, FAR, stage, cursor);
    else
        k_panel<RPT, NBW, false><<<1, PT, shm, ctx->stream>>>(W, ldw, m, kr0, c0, w, mp, ple, invtab, ipiv, pivval,
                                                             pivinv, pivcol, colmap, map_dst, map_src, ctl, abort,
                                                             KC, FAR, stage, cursor);
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

__global__ void k_set_int(int *p, int v) { *p = v; }

__global__ void k_inv_table(uint32_t *__restrict__ t, ModP mp)
{
    for (int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; a < mp.p; a += (int64_t)gridDim.x * blockDim.x)
        t[a] = a == 0 ? 0u : modinv((uint32_t)a, mp);
}

// Inverses of every residue for p <= 2^22 (at most 16 MB), built once per
// modulus and kept in the context: the panel's pivot inverse becomes one load.
const uint32_t *inverse_table(modp_ctx *ctx, const ModP &mp, int *st)
{
    if (mp.p > ((int64_t)1 << 22)) return nullptr;
    if (ctx->invtab && ctx->invtab_p == mp.p) return ctx->invtab;
    const size_t need = sizeof(uint32_t) * (size_t)mp.p;
    if (ctx->invtab_cap < need) {
        if (ctx->invtab) {
            hipError_t e = hipStreamSynchronize(ctx->stream);
            const hipError_t ef = hipFree(ctx->invtab);
            ctx->invtab = nullptr;
            ctx->invtab_cap = 0;
            if (e == hipSuccess) e = ef;
            if (e != hipSuccess) { *st = hip_check(ctx, e, "inverse table release"); return nullptr; }
        }
        hipError_t e = hipMalloc(&ctx->invtab, need);
        if (e != hipSuccess) { *st = hip_check(ctx, e, "inverse table"); return nullptr; }
        ctx->invtab_cap = need;
    }
    k_inv_table<<<grid_for(mp.p, 4096), 256, 0, ctx->stream>>>(ctx->invtab, mp);
    if (hipGetLastError() != hipSuccess) { *st = MODP_ERR_DEVICE; return nullptr; }
    ctx->invtab_p = mp.p;
    return ctx->invtab;
}


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
    if (m > (int64_t)1 << 30)
        return set_error(ctx, MODP_ERR_DIM, "elimination supports at most 2^30 rows (got %lld)", (long long)m);

    // Panel shape of an outer block starting at row KR: the fewest rows per
    // thread that hold every row from KR's 4-aligned row down (the panel
    // kernel reads 4-row vectors from kr0 & ~3 >= KR & ~3), and the panel
    // width for that many rows per thread (rows per thread x width <= 64
    // registers; widths measured at 2048 and 4096 rows).  Late blocks have
    // fewer rows, so their panels are shorter and wider.
    const char *wenv = getenv("MODP_PANEL_W");
    const int want_w = wenv ? atoi(wenv) : 0;
    // MODP_PANEL_GLOBAL=1 sends every panel to the W-resident kernel (used
    // to compare the two panel kernels at sizes both handle)
    const char *genv = getenv("MODP_PANEL_GLOBAL");
    const bool force_global = genv && atoi(genv) != 0;
    auto panel_shape = [&](int64_t KR, int *RPT, int *NBW) {
        const int64_t need = m - (KR & ~(int64_t)3);
        if (force_global || need > (int64_t)32 * PT) {
            // more rows than the register panel holds: W-resident panel
            *RPT = 0;
            *NBW = GNBW;
            return;
        }
        int r = 1;
        while ((int64_t)r * PT < need) r *= 2;
        int wd = r == 1 ? 32 : (r == 2 ? 16 : (r <= 8 ? 8 : 2));
        // tuning override among the instantiated panel widths
        if ((r <= 2 && (want_w == 16 || want_w == 32)) || (r == 4 && (want_w == 8 || want_w == 16)) ||
            (r == 8 && (want_w == 4 || want_w == 8)))
            wd = want_w;
        *RPT = r;
        *NBW = wd;
    };

    int st = MODP_OK;
    const int64_t ldw = rup(m, 64);
    int32_t *W = (int32_t *)ws_get(ctx, WS_WORK, sizeof(int32_t) * (size_t)ldw * (size_t)n, &st);
    if (!W) return st;
    // pivot arrays, per-panel control blocks and row gather maps of one outer block
    const size_t npv = (size_t)mn + 64;
    const size_t piv_bytes = sizeof(int32_t) * (4 * npv + 64 + 2 * MAXPAN * 64) + sizeof(PanelCtl) * MAXPAN + 256 +
                             sizeof(int32_t) * 64 * NBO;
    int32_t *piv = (int32_t *)ws_get(ctx, WS_PIV, piv_bytes, &st);
    if (!piv) return st;
    int32_t *ipiv = piv, *pivval = piv + npv, *pivinv = piv + 2 * npv, *pivcol = piv + 3 * npv;
    int32_t *colmap = piv + 4 * npv;
    int32_t *bmap_dst = colmap + 64, *bmap_src = bmap_dst + MAXPAN * 64;
    PanelCtl *bctl = (PanelCtl *)(bmap_src + MAXPAN * 64);
    int *flag = (int *)(bctl + MAXPAN);
    int32_t *stage = (int32_t *)(flag + 64);  // moved rows' old values, 64 x NBO
    int *cursor = flag + 2;                   // first row of the next panel (lu, rref)
    int32_t *Ubig = (int32_t *)ws_get(ctx, WS_UBUF, sizeof(int32_t) * (size_t)NBO * (size_t)(n + 1), &st);
    if (!Ubig) return st;
    // row positions of the W-resident panel (position of each row, row at
    // each position)
    int32_t *gpos = nullptr, *grow = nullptr;
    if (force_global || m > (int64_t)32 * PT) {
        gpos = (int32_t *)ws_get(ctx, WS_GPOS, sizeof(int32_t) * 2 * (size_t)m, &st);
        if (!gpos) return st;
        grow = gpos + m;
    }
    PanelCtl *hctl = (PanelCtl *)pinned_get(ctx, sizeof(PanelCtl) * (MAXPAN + 1) + 64, &st);
    if (!hctl) return st;
    int *abort_flag = flag + 1;

    MODP_TRY(util_convert(ctx, kind, m, n, A, lda, EK_I32, W, ldw, mp));
    const uint32_t *invtab = inverse_table(ctx, mp, &st);
    if (st != MODP_OK) return st;

    auto gemm_sub = [&](int64_t r0, int64_t rows, int64_t ca, int64_t cols, int64_t kcol, int64_t K,
                        const int32_t *Ub, int use_form, const int *skip) -> int {
        if (rows <= 0 || cols <= 0 || K <= 0) return MODP_OK;
        GemmArgs g;
        g.m = rows; g.n = cols; g.k = K;
        g.A = W + r0 + kcol * ldw; g.lda = ldw; g.ka = EK_I32;
        g.B = Ub; g.ldb = NBO; g.kb = EK_I32;
        g.C = W + r0 + ca * ldw; g.ldc = ldw; g.kc = EK_I32;
        g.epi = EPI_SUB;
        g.mp = mp;
        g.skip = skip;
        if (use_form == MODP_FORM_LIMB_I8) return gemm_limb_i8(ctx, g);
        return gemm_plain_f64(ctx, g);
    };
    // the trailing updates have K = 128, below the general limb threshold,
    // but at 1024 x 1024 and beyond the limb form's split passes are paid
    // back (4096^2 PLUQ: 17.6 ms limb vs 19.0 ms plain)
    const int big_form = form != MODP_FORM_AUTO ? form
                         : (m >= 1024 && n >= 1024) ? MODP_FORM_LIMB_I8 : choose_form(m, n, NBO, p);

    // panel t of the current outer block at (rj, cj), width w
    auto launch_panel_t = [&](int RPT, int NBW, int t, int64_t KC, int64_t FAR, int64_t rj, int64_t cj, int64_t w,
                              int *abort, int *cursor) -> int {
        if (t >= MAXPAN) return set_error(ctx, MODP_ERR_DEVICE, "more than %d panels in an outer block", MAXPAN);
        int32_t *mdst = bmap_dst + t * 64, *msrc = bmap_src + t * 64;
        PanelCtl *pctl = bctl + t;
        if (RPT == 0) {
            TimedLaunch tl(ctx, "elim_panel");
            k_panel_g<GNBW><<<1, GPT, 0, ctx->stream>>>(W, ldw, (int)m, (int)rj, (int)cj, (int)w, mp, ple, invtab, ipiv,
                                                        pivval, pivinv, pivcol, colmap, mdst, msrc, pctl, abort,
                                                        (int)KC, (int)FAR, stage, cursor, gpos, grow);
            return hip_check(ctx, hipGetLastError(), "k_panel_g launch");
        }
        const int key = RPT * 100 + NBW;
#define LAUNCH_PANEL(R_, W_) \
    case R_ * 100 + W_: \
        return launch_panel<R_, W_>(ctx, W, ldw, (int)m, (int)rj, (int)cj, (int)w, mp, ple, invtab, ipiv, pivval, \
                                    pivinv, pivcol, colmap, mdst, msrc, pctl, abort, (int)KC, (int)FAR, stage, cursor);
        switch (key) {
            LAUNCH_PANEL(1, 16)
            LAUNCH_PANEL(1, 32)
            LAUNCH_PANEL(2, 16)
            LAUNCH_PANEL(2, 32)
            LAUNCH_PANEL(4, 8)
            LAUNCH_PANEL(4, 16)
            LAUNCH_PANEL(8, 4)
            LAUNCH_PANEL(8, 8)
            LAUNCH_PANEL(16, 2)
            LAUNCH_PANEL(32, 2)
        default:
            return set_error(ctx, MODP_ERR_ARG, "no panel kernel for %d rows/thread x %d columns", RPT, NBW);
        }
#undef LAUNCH_PANEL
    };
    // the rest of the outer block after panel t: its row moves, and right
    // of the panel the rows of U and the rank-npiv update below the pivots
    auto update_t = [&](int t, int64_t KC, int64_t FAR, int64_t rj, int64_t cj, int64_t w, const int *skip) -> int {
        const int64_t nl = cj - KC, nr = FAR - (cj + w);
        const unsigned gy = (unsigned)((nl + UT_C - 1) / UT_C + (nr + UT_C - 1) / UT_C);
        if (gy == 0) return MODP_OK;
        const dim3 grid((unsigned)((m - rj + UT_R - 1) / UT_R), gy);
        TimedLaunch tl(ctx, "elim_update");
#define LAUNCH_UPDATE(WM_) \
    if (mp.small) \
        k_update<WM_, true><<<grid, 256, 0, ctx->stream>>>(W, ldw, (int)m, (int)KC, (int)FAR, (int)cj, (int)w, \
                                                           bctl + t, pivinv, pivcol, bmap_dst + t * 64, stage, mp, skip); \
    else \
        k_update<WM_, false><<<grid, 256, 0, ctx->stream>>>(W, ldw, (int)m, (int)KC, (int)FAR, (int)cj, (int)w, \
                                                            bctl + t, pivinv, pivcol, bmap_dst + t * 64, stage, mp, skip);
        if (w <= 8) {
            LAUNCH_UPDATE(8)
        } else if (w <= 16) {
            LAUNCH_UPDATE(16)
        } else {
            LAUNCH_UPDATE(32)
        }
#undef LAUNCH_UPDATE
        return hip_check(ctx, hipGetLastError(), "k_update launch");
    };
    // end of an outer block with P pivots over `consumed` columns
    auto block_end = [&](int64_t KR, int64_t KC, int64_t FAR, int64_t P, int64_t consumed, int npanels,
                         const int *skip) -> int {
        if (P <= 0) return MODP_OK;
        const int64_t nout = n - (FAR - KC);
        if (nout > 0) {
            TimedLaunch tl(ctx, "elim_block_finish");
            int64_t g = (nout + 15) / 16;
            if (g > 2048) g = 2048;
            if (mp.small)
                k_block_finish<true><<<(unsigned)g, 1024, 0, ctx->stream>>>(
                    W, ldw, n, (int)KR, (int)KC, (int)FAR, (int)P, (int)consumed, npanels, bmap_dst, bmap_src, bctl,
                    pivinv, pivcol, Ubig, NBO, mp, skip);
            else
                k_block_finish<false><<<(unsigned)g, 1024, 0, ctx->stream>>>(
                    W, ldw, n, (int)KR, (int)KC, (int)FAR, (int)P, (int)consumed, npanels, bmap_dst, bmap_src, bctl,
                    pivinv, pivcol, Ubig, NBO, mp, skip);
            HIP_TRY(ctx, hipGetLastError());
        }
        const int64_t rbeg = KR + P;
        return gemm_sub(rbeg, m - rbeg, FAR, n - FAR, KC, consumed, Ubig + (FAR - KC) * NBO, big_form, skip);
    };

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
        int RPT, NBW;
        panel_shape(KR, &RPT, &NBW);
        bool stop_zero = false;
        if (!ple && nbo > 0) {
            // pluq: every panel of the block is enqueued assuming it finds a
            // pivot in each column; a panel that meets a zero column raises
            // the abort flag, which turns the rest of the block into no-ops,
            // and the block is completed from the host after one sync
            HIP_TRY(ctx, hipMemsetAsync(abort_flag, 0, sizeof(int), ctx->stream));
            std::vector<int64_t> pr, pcj, pw;
            for (int64_t cj = KC; cj < FAR;) {
                const int64_t rj = KR + (cj - KC);
                if (rj >= m) break;
                int64_t w = FAR - cj;
                if (w > NBW) w = NBW;
                if (w > m - rj) w = m - rj;
                pr.push_back(rj); pcj.push_back(cj); pw.push_back(w);
                cj += w;
            }
            const int np = (int)pr.size();
            int64_t Pfull = 0;
            for (int t = 0; t < np; ++t) {
                MODP_TRY(launch_panel_t(RPT, NBW, t, KC, FAR, pr[t], pcj[t], pw[t], abort_flag, nullptr));
                MODP_TRY(update_t(t, KC, FAR, pr[t], pcj[t], pw[t], abort_flag));
                Pfull += pw[t];
            }
            MODP_TRY(block_end(KR, KC, FAR, Pfull, Pfull, np, abort_flag));
            // the control blocks and, right after them, the flags: one copy
            HIP_TRY(ctx, hipMemcpyAsync(hctl, bctl, sizeof(PanelCtl) * MAXPAN + 2 * sizeof(int), hipMemcpyDeviceToHost,
                                        ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            const int aborted = reinterpret_cast<const int *>(hctl + MAXPAN)[1];
            if (!aborted) {
                kr += Pfull;
                kc += Pfull;
            } else {
                int ts = 0;
                while (ts < np && hctl[ts].status == 0) ++ts;
                if (ts >= np)
                    return set_error(ctx, MODP_ERR_DEVICE, "elimination pipeline aborted without a stopping panel");
                if (hctl[ts].status == 3)
                    return set_error(ctx, MODP_ERR_NOT_INVERTIBLE,
                                     "pivot has no inverse modulo %lld (modulus not prime)", (long long)p);
                const int64_t npiv = hctl[ts].npiv;
                int64_t P = 0;
                for (int t = 0; t < ts; ++t) P += pw[t];
                if (npiv > 0) MODP_TRY(update_t(ts, KC, FAR, pr[ts], pcj[ts], pw[ts], nullptr));
                P += npiv;
                MODP_TRY(block_end(KR, KC, FAR, P, P, ts + 1, nullptr));
                kr += P;
                kc += P;
                stop_zero = hctl[ts].status == 1;
            }
        } else if (nbo > 0) {
            // ple: panels skip zero columns, so a panel's first row depends
            // on the pivots found before it.  The panels of the block are
            // enqueued at their fixed columns and each takes its first row
            // from a device cursor that the previous one advanced; the
            // updates read it from the panel's control block.  A panel that
            // runs out of rows (or meets a pivot without inverse) raises the
            // abort flag, which turns the rest of the block into no-ops, and
            // the host reads the outcome once per block.
            HIP_TRY(ctx, hipMemsetAsync(abort_flag, 0, sizeof(int), ctx->stream));
            k_set_int<<<1, 1, 0, ctx->stream>>>(cursor, (int)KR);
            HIP_TRY(ctx, hipGetLastError());
            std::vector<int64_t> pcj, pw;
            for (int64_t cj = KC; cj < FAR; cj += NBW) {
                pcj.push_back(cj);
                pw.push_back(std::min<int64_t>(NBW, FAR - cj));
            }
            const int np = (int)pcj.size();
            for (int t = 0; t < np; ++t) {
                MODP_TRY(launch_panel_t(RPT, NBW, t, KC, FAR, KR, pcj[t], pw[t], abort_flag, cursor));
                MODP_TRY(update_t(t, KC, FAR, KR, pcj[t], pw[t], abort_flag));
            }
            HIP_TRY(ctx, hipMemcpyAsync(hctl, bctl, sizeof(PanelCtl) * np, hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            int64_t P = 0, consumed = 0;
            int npanels = 0;
            for (int t = 0; t < np; ++t) {
                const PanelCtl &pc = hctl[t];
                if (pc.status == 4) break;
                ++npanels;
                if (pc.status == 3)
                    return set_error(ctx, MODP_ERR_NOT_INVERTIBLE,
                                     "pivot has no inverse modulo %lld (modulus not prime)", (long long)p);
                P += pc.npiv;
                consumed += pc.ncols;
                if (pc.status != 0) {
                    // the update after this panel was skipped by its own flag
                    if (pc.npiv > 0) MODP_TRY(update_t(t, KC, FAR, KR, pcj[t], pw[t], nullptr));
                    break;
                }
            }
            kr += P;
            kc += consumed;
            MODP_TRY(block_end(KR, KC, FAR, P, consumed, npanels, nullptr));
            if (P < consumed && kr < m && kc < n) {
                // the block met zero columns: if nothing nonzero is left
                // below row kr, no later column has a pivot and the
                // factorisation is complete (rank = kr)
                HIP_TRY(ctx, hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
                k_any_nonzero<<<grid_for((m - kr) * (n - kc), 4096), 256, 0, ctx->stream>>>(W, ldw, kr, m, kc, n, flag);
                HIP_TRY(ctx, hipGetLastError());
                int *hflag = (int *)hctl;
                HIP_TRY(ctx, hipMemcpyAsync(hflag, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
                HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
                if (!*hflag) done = true;
            }
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
        const int bform = form != MODP_FORM_AUTO ? form
                          : (m >= 1024 && n >= 1024) ? MODP_FORM_LIMB_I8 : MODP_FORM_PLAIN_F64;
        for (int64_t i1 = rank; i1 > 0;) {
            const int64_t i0 = i1 > 64 ? i1 - 64 : 0;
            const int nb = (int)(i1 - i0);
            // first pivot column of the block (the factorisation returned
            // the pivot columns, 1-based)
            const int64_t cs = res.pivcols[(size_t)i0] - 1;
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
