"""Multi-modular plumbing on the GPU: Chinese remaindering of per-prime
results into integers (against Python integers), the multi-modular integer
product end to end, and the modulus helpers mod_elements! / change_modulus /
change_modulus_no_alloc! (CuModMatrix.jl:712-758)."""
import math

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def crt_py(res, moduli, signed):
    """Reference reconstruction in Python integers (iterative CRT)."""
    y, M = 0, 1
    for x, q in zip(res, moduli):
        t = ((int(x) - y) * pow(M, -1, q)) % q
        y += M * t
        M *= q
    if signed and y > M // 2:
        y -= M
    return y


MODULI = {
    "one": [65521],
    "two_small": [2, 3],
    "mixed": [2, 3, 7, 251, 65521, 2 ** 26 - 5],
    "twelve_large": O.largest_primes_below(2 ** 26, 12),
    "composite_coprime": [2 ** 26, 3 ** 16, 5 ** 11, 7 ** 9],
}


@pytest.mark.parametrize("name", sorted(MODULI))
@pytest.mark.parametrize("signed", [False, True])
def test_crt_matches_python_integers(gffm, name, signed):
    g = gffm
    moduli = MODULI[name]
    m, n = 37, 29
    rng = np.random.default_rng(len(moduli))
    res = [rng.integers(0, q, size=(m, n)) for q in moduli]
    # extremes: all zero, all q-1 (y = P-1, i.e. -1 signed), and P/2 region
    for R, q in zip(res, moduli):
        R[0, 0] = 0
        R[1, 0] = q - 1
    mats = [g.CuModMatrix(R, q, elem_type=np.float64) for R, q in zip(res, moduli)]
    Y = g.crt_reconstruct(mats, signed=signed)
    P = math.prod(moduli)
    for i in range(m):
        for j in range(n):
            want = crt_py([R[i, j] for R in res], moduli, signed)
            assert Y[i, j] == want, (i, j)
    assert Y[0, 0] == 0 and Y[1, 0] == (-1 if signed else P - 1)


def test_crt_half_boundary_signed(gffm):
    # y = floor(P/2) stays positive, floor(P/2) + 1 becomes negative
    g = gffm
    moduli = [251, 65521, 2 ** 26 - 5]
    P = math.prod(moduli)
    vals = [P // 2, P // 2 + 1, 1, P - 1]
    res = [np.array([[v % q for v in vals]]) for q in moduli]
    mats = [g.CuModMatrix(R, q, elem_type=np.float64) for R, q in zip(res, moduli)]
    Y = g.crt_reconstruct(mats, signed=True)
    assert list(Y[0]) == [P // 2, P // 2 + 1 - P, 1, -1]


def test_multimodular_integer_product(gffm):
    # C = A B over the integers (|entries| < 2^40, k = 96, |C| < 2^87): the
    # product under four primes near 2^26 in one batched launch, then signed
    # CRT; equal to the exact Python-integer product
    g = gffm
    rng = np.random.default_rng(5)
    m, k, n = 45, 96, 38
    Ah = rng.integers(-(2 ** 40) + 1, 2 ** 40, size=(m, k), dtype=np.int64)
    Bh = rng.integers(-(2 ** 40) + 1, 2 ** 40, size=(k, n), dtype=np.int64)
    primes = O.largest_primes_below(2 ** 26, 4)
    assert math.prod(primes) > 2 * k * (2 ** 40) ** 2
    jobs, outs = [], []
    for q in primes:
        A = g.CuModMatrix(np.mod(Ah, q), q, elem_type=np.float64)
        B = g.CuModMatrix(np.mod(Bh, q), q, elem_type=np.float64)
        C = g.zeros(np.float64, m, n, q)
        jobs.append((A, B, C, q))
        outs.append(C)
    g.gemm_batched(jobs)
    Y = g.crt_reconstruct(outs, signed=True)
    exact = Ah.astype(object).dot(Bh.astype(object))
    assert (Y == exact).all()


def test_crt_errors(gffm):
    g = gffm
    A = g.CuModMatrix(np.ones((3, 3), dtype=np.int64), 6, elem_type=np.float64)
    B = g.CuModMatrix(np.ones((3, 3), dtype=np.int64), 9, elem_type=np.float64)
    with pytest.raises(g.ArgumentError):
        g.crt_reconstruct([A, B])  # gcd(6, 9) = 3
    C = g.CuModMatrix(np.ones((2, 3), dtype=np.int64), 7, elem_type=np.float64)
    with pytest.raises(g.CuModArraySizeMismatchException):
        g.crt_reconstruct([A, C])


@pytest.mark.parametrize("et", [np.float64, np.float32])
def test_modulus_helpers(gffm, et):
    g = gffm
    rng = np.random.default_rng(6)
    Ah = rng.integers(0, 1000, size=(40, 30))
    A = g.CuModMatrix(Ah, 1009, elem_type=et)
    # mod_elements!(A, 7): every stored entry, padding included, mod 7
    pad = A.padded_array().copy()
    g.mod_elements_(A, 7)
    assert np.array_equal(A.padded_array(), np.mod(pad, 7))
    # change_modulus to a smaller modulus reduces, to a larger one copies
    A = g.CuModMatrix(Ah, 1009, elem_type=et)
    R = g.change_modulus(A, 13)
    assert R.N == 13 and np.array_equal(g.Array(R).astype(np.int64), np.mod(Ah, 13))
    assert np.array_equal(g.Array(A).astype(np.int64), Ah)  # A untouched
    R2 = g.change_modulus(A, 65521)
    assert R2.N == 65521 and np.array_equal(g.Array(R2).astype(np.int64), Ah)
    # change_modulus_no_alloc!: in place
    S = g.change_modulus_no_alloc_(A, 11)
    assert S.N == 11 and S.data is A.data
    assert np.array_equal(g.Array(A).astype(np.int64), np.mod(Ah, 11))
