// crt.hip -- Chinese remaindering of r residue matrices (r <= 12 pairwise
// coprime moduli <= 2^26) into multi-word integers: the per-prime results of
// the batched forms turned back into one integer matrix (the multi-modular
// pattern; the reference's nearest counterpart is the two-limb composite
// modulus of src/KaratsubaMatrix/KaratsubaMatrix.jl:18-51).
//
// Per entry, Garner's algorithm (Garner 1959; Knuth TAOCP vol. 2, 4.3.2,
// algorithm for mixed-radix conversion) gives the mixed-radix digits
// y = d_0 + d_1 p_0 + d_2 p_0 p_1 + ..., 0 <= d_i < p_i, with
// d_i = ((x_i - d_0) c_0i - d_1) c_1i - ... - d_{i-1}) c_{i-1,i} mod p_i and
// c_ji = p_j^{-1} mod p_i; Horner's rule then turns the digits into the
// binary value in 32-bit words, y in [0, P) for P = prod p_i, or, signed,
// the representative in (-P/2, P/2] in two's complement.
#include "common.h"

namespace {

constexpr int CRT_MAX = 12;
constexpr int CRT_WORDS = 10;  // 12 moduli of 26 bits and a sign bit: 313 bits

struct CrtConsts {
    int r;
    int sign;
    ModP mp[CRT_MAX];
    uint32_t c[CRT_MAX][CRT_MAX];  // c[j][i] = p_j^{-1} mod p_i for j < i
    uint32_t P[CRT_WORDS];         // prod p_i, little-endian words
    uint32_t H[CRT_WORDS];         // floor(P / 2)
};

struct CrtInputs {
    const void *X[CRT_MAX];
    int64_t ld[CRT_MAX];
};

template <typename T>
__global__ __launch_bounds__(256) void k_crt(int64_t rows, int64_t cols, CrtInputs in, uint32_t *__restrict__ Y,
                                             int64_t ldy, int64_t plane, int words, CrtConsts cc)
{
    const int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % rows, j = e / rows;
        uint32_t d[CRT_MAX];
#pragma unroll
        for (int a = 0; a < CRT_MAX; ++a) {
            d[a] = 0;
            if (a >= cc.r) continue;
            const ModP &pa = cc.mp[a];
            uint32_t t = to_res<T>(((const T *)in.X[a])[i + j * in.ld[a]], pa);
#pragma unroll
            for (int b = 0; b < CRT_MAX; ++b) {
                if (b >= a) break;
                const uint32_t db = (uint32_t)dmod((double)d[b], pa);
                t = mulmod(submod(t, db, pa), cc.c[b][a], pa);
            }
            d[a] = t;
        }
        // Horner: y = d_{r-1}; y = y p_a + d_a for a = r-2 .. 0
        uint32_t y[CRT_WORDS];
#pragma unroll
        for (int w = 0; w < CRT_WORDS; ++w) y[w] = 0;
#pragma unroll
        for (int a = CRT_MAX - 1; a >= 0; --a) {
            if (a >= cc.r) continue;
            if (a == cc.r - 1) {
                y[0] = d[a];
                continue;
            }
            uint64_t carry = d[a];
            const uint64_t pa = cc.mp[a].p32;
#pragma unroll
            for (int w = 0; w < CRT_WORDS; ++w) {
                const uint64_t t = (uint64_t)y[w] * pa + carry;
                y[w] = (uint32_t)t;
                carry = t >> 32;
            }
        }
        if (cc.sign) {
            // y > floor(P/2): take y - P (two's complement)
            bool gt = false, decided = false;
#pragma unroll
            for (int w = CRT_WORDS - 1; w >= 0; --w) {
                if (!decided && y[w] != cc.H[w]) {
                    gt = y[w] > cc.H[w];
                    decided = true;
                }
            }
            if (gt) {
                uint64_t borrow = 0;
#pragma unroll
                for (int w = 0; w < CRT_WORDS; ++w) {
                    const uint64_t t = (uint64_t)y[w] - cc.P[w] - borrow;
                    y[w] = (uint32_t)t;
                    borrow = (t >> 32) & 1u;
                }
            }
        }
#pragma unroll
        for (int w = 0; w < CRT_WORDS; ++w)
            if (w < words) Y[w * plane + i + j * ldy] = y[w];
    }
}

int64_t gcd64(int64_t a, int64_t b)
{
    while (b) {
        const int64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

int64_t inv_mod(int64_t a, int64_t m)
{
    // extended Euclid (Knuth TAOCP vol. 2, 4.5.2)
    int64_t t = 0, nt = 1, r = m, nr = a % m;
    while (nr != 0) {
        const int64_t q = r / nr;
        int64_t x = t - q * nt; t = nt; nt = x;
        x = r - q * nr; r = nr; nr = x;
    }
    return t < 0 ? t + m : t;
}

}  // namespace

// words needed for the product of the moduli (plus a sign bit when signed)
int crt_words(const int64_t *p, int r, int sign)
{
    double bits = sign ? 1.0 : 0.0;
    for (int i = 0; i < r; ++i) bits += log2((double)p[i]);
    return (int)ceil((bits + 1e-9) / 32.0);
}

int util_crt(modp_ctx *ctx, int kind, int64_t rows, int64_t cols, int r, const void *const *X, const int64_t *ldx,
             const int64_t *p, int sign, uint32_t *Y, int64_t ldy, int64_t words)
{
    if (r < 1 || r > CRT_MAX) return set_error(ctx, MODP_ERR_ARG, "CRT of %d moduli (1..%d supported)", r, CRT_MAX);
    for (int a = 0; a < r; ++a) {
        if (p[a] < 2 || p[a] > ((int64_t)1 << 26))
            return set_error(ctx, MODP_ERR_MODULUS_RANGE, "modulus %lld outside 2..2^26", (long long)p[a]);
        for (int b = 0; b < a; ++b)
            if (gcd64(p[a], p[b]) != 1)
                return set_error(ctx, MODP_ERR_MODULUS_RANGE, "moduli %lld and %lld are not coprime",
                                 (long long)p[b], (long long)p[a]);
    }
    const int need = crt_words(p, r, sign);
    if (words < need || words > CRT_WORDS)
        return set_error(ctx, MODP_ERR_ARG, "CRT output needs %d words of 32 bits, %lld given", need, (long long)words);
    if (rows <= 0 || cols <= 0) return MODP_OK;
    CrtConsts cc{};
    cc.r = r;
    cc.sign = sign ? 1 : 0;
    CrtInputs in{};
    for (int a = 0; a < r; ++a) {
        cc.mp[a] = make_modp(p[a]);
        in.X[a] = X[a];
        in.ld[a] = ldx[a];
        for (int b = 0; b < a; ++b) cc.c[b][a] = (uint32_t)inv_mod(p[b] % p[a], p[a]);
    }
    // P = prod p_i in words (schoolbook), H = P >> 1
    uint32_t P[CRT_WORDS] = {1};
    for (int a = 0; a < r; ++a) {
        uint64_t carry = 0;
        for (int w = 0; w < CRT_WORDS; ++w) {
            const uint64_t t = (uint64_t)P[w] * (uint64_t)p[a] + carry;
            P[w] = (uint32_t)t;
            carry = t >> 32;
        }
    }
    for (int w = 0; w < CRT_WORDS; ++w) {
        cc.P[w] = P[w];
        cc.H[w] = (P[w] >> 1) | (w + 1 < CRT_WORDS ? (P[w + 1] & 1u) << 31 : 0u);
    }
    const int64_t total = rows * cols;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    const int64_t plane = ldy * cols;
    if (kind == EK_F64)
        k_crt<double><<<(unsigned)blocks, 256, 0, ctx->stream>>>(rows, cols, in, Y, ldy, plane, (int)words, cc);
    else if (kind == EK_F32)
        k_crt<float><<<(unsigned)blocks, 256, 0, ctx->stream>>>(rows, cols, in, Y, ldy, plane, (int)words, cc);
    else
        return set_error(ctx, MODP_ERR_DTYPE, "crt: unsupported element type");
    HIP_TRY(ctx, hipGetLastError());
    return MODP_OK;
}
