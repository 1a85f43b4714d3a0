"""GEMM parity on the GPU through the C ABI: bit-exact against the oracle."""
import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))
FORMS = {"plain": 1, "limb": 2}


def dev(g, X, p, et=np.float64):
    return g.CuModMatrix(np.asarray(X), p, elem_type=et)


@pytest.mark.parametrize("form", ["plain", "limb"])
@pytest.mark.parametrize("et", [np.float64, np.float32])
@pytest.mark.parametrize("case", GOLD["matmul"], ids=lambda c: c["name"])
def test_known_answers(gffm, case, et, form):
    g = gffm
    A = dev(g, case["A"], case["p"], et)
    B = dev(g, case["B"], case["p"], et)
    C = g.zeros(et, len(case["A"]), len(case["B"][0]), case["p"])
    g.mul_(C, A, B, form=FORMS[form])
    assert g.Array(C).astype(np.int64).tolist() == case["C"]
    assert (A * B).N == case["p"]
    assert g.Array(A * B).astype(np.int64).tolist() == case["C"]


def test_power_and_mat_mul_gpu_type(gffm):
    g = gffm
    G = dev(g, [[1, 2, 3], [0, 1, 4], [5, 6, 0]], 11)
    assert g.Array(G ** 2).astype(int).tolist() == [[5, 0, 0], [9, 3, 4], [5, 5, 6]]
    assert g.Array(G ** 0).astype(int).tolist() == np.eye(3, dtype=int).tolist()
    A = dev(g, [[1, 2, 3], [4, 5, 6]], 11)
    B = dev(g, [[7, 8], [9, 10], [0, 1]], 11)
    assert g.Array(g.mat_mul_gpu_type(A, B, 7)).astype(int).tolist() == [[4, 3], [3, 4]]


@pytest.mark.parametrize("form", ["plain", "limb"])
def test_config1_1024_p65521_seed0(gffm, form):
    g = gffm
    p, n = 65521, 1024
    Ah = O.draw_matrix(n, n, p, seed=0)
    Bh = O.draw_matrix(n, n, p, seed=0, offset=n * n)
    A, B = dev(g, Ah, p), dev(g, Bh, p)
    C = g.zeros(np.float64, n, n, p)
    g.mul_(C, A, B, form=FORMS[form])
    ref = O.gemm(Ah, Bh, p)
    got = g.Array(C).astype(np.int64)
    assert int(np.abs(got - ref).max()) == 0
    # the padding stays zero
    assert not g.Array(C).dtype == object
    full = C.padded_array()
    assert not full[n:, :].any() and not full[:, n:].any()


SHAPES = [(1, 1, 1), (31, 32, 33), (33, 31, 32), (64, 64, 64), (100, 1, 257), (1, 100, 5), (257, 130, 65),
          (129, 257, 1000)]
PRIMES = [2, 3, 127, 251, 65521, 8388593, 2 ** 26 - 5]


@pytest.mark.parametrize("form", ["plain", "limb"])
@pytest.mark.parametrize("p", PRIMES)
@pytest.mark.parametrize("shape", SHAPES, ids=str)
def test_shapes_and_primes(gffm, shape, p, form):
    g = gffm
    m, n, k = shape
    print(f"gemm {form} shape={shape} p={p}", flush=True)
    Ah = O.draw_matrix(m, k, p, seed=7)
    Bh = O.draw_matrix(k, n, p, seed=7, offset=m * k)
    C = g.zeros(np.float64, m, n, p)
    g.mul_(C, dev(g, Ah, p), dev(g, Bh, p), form=FORMS[form])
    assert np.array_equal(g.Array(C).astype(np.int64), O.gemm(Ah, Bh, p))


@pytest.mark.parametrize("form", ["plain", "limb"])
@pytest.mark.parametrize("p", [2 ** 26 - 5, 2 ** 26, 65521, 127, 2])
def test_extreme_residues(gffm, p, form):
    # all entries p-1: every overflow bound at its limit
    g = gffm
    m, n, k = 70, 90, 3000
    Ah = np.full((m, k), p - 1, dtype=np.int64)
    Bh = np.full((k, n), p - 1, dtype=np.int64)
    C = g.zeros(np.float64, m, n, p)
    g.mul_(C, dev(g, Ah, p), dev(g, Bh, p), form=FORMS[form])
    assert np.all(g.Array(C).astype(np.int64) == k % p)


@pytest.mark.parametrize("p", [257, 65537, 16777213, 16777259])
def test_limb_digit_count_boundaries(gffm, p):
    # moduli just above and below 256^L: the signed base-256 digit count
    # changes there (257 -> 2 digits, 65537 -> 3, 2^24-3 -> 3, 2^24+43 -> 4)
    g = gffm
    m, n, k = 65, 70, 300
    print(f"gemm limb digit boundary p={p}", flush=True)
    Ah = O.draw_matrix(m, k, p, seed=11)
    Bh = O.draw_matrix(k, n, p, seed=11, offset=m * k)
    Ah[0, :] = p - 1
    Bh[:, 0] = p - 1
    C = g.zeros(np.float64, m, n, p)
    g.mul_(C, dev(g, Ah, p), dev(g, Bh, p), form=2)
    assert np.array_equal(g.Array(C).astype(np.int64), O.gemm(Ah, Bh, p))


def test_limb_extreme_digits_two_chunks(gffm):
    # p = 65521: the residue 32640 is represented as 32640 - p = -32881,
    # digits (-113, -128), the largest digit products the 2-digit form meets;
    # K = 70,000 spans two int32-safe chunks (65,472 each)
    g = gffm
    p, m, n, k = 65521, 33, 17, 70000
    print("gemm limb extreme digits p=65521 k=70000", flush=True)
    Ah = np.full((m, k), 32640, dtype=np.int64)
    Bh = np.full((k, n), 32640, dtype=np.int64)
    Ah[1, :] = O.draw_matrix(1, k, p, seed=5)[0]
    C = g.zeros(np.float64, m, n, p)
    g.mul_(C, dev(g, Ah, p), dev(g, Bh, p), form=2)
    assert np.array_equal(g.Array(C).astype(np.int64), O.gemm(Ah, Bh, p))


def test_limb_k_chunking(gffm):
    # K beyond one int32-safe chunk of the 4-digit form (32,704): chunks are
    # recombined mod p
    g = gffm
    p, m, n, k = 2 ** 26 - 5, 40, 24, 70000
    Ah = O.draw_matrix(m, k, p, seed=3)
    Bh = O.draw_matrix(k, n, p, seed=3, offset=m * k)
    Ah[:, :5000] = p - 1
    Bh[:5000, :] = p - 1
    C = g.zeros(np.float64, m, n, p)
    g.mul_(C, dev(g, Ah, p), dev(g, Bh, p), form=2)
    assert np.array_equal(g.Array(C).astype(np.int64), O.gemm(Ah, Bh, p))


@pytest.mark.parametrize("form", ["plain", "limb"])
def test_float32_storage(gffm, form):
    g = gffm
    for p in (2 ** 11, 11 ** 3, 2897):
        Ah = O.draw_matrix(100, 100, p, seed=1)
        Bh = O.draw_matrix(100, 100, p, seed=2)
        Ah[4, 4] = p - 5
        Bh[94, 94] = p - 23
        C = g.zeros(np.float32, 100, 100, p)
        g.mul_(C, dev(g, Ah, p, np.float32), dev(g, Bh, p, np.float32), form=FORMS[form])
        assert np.array_equal(g.Array(C).astype(np.int64), O.gemm(Ah, Bh, p))


def test_float32_range_error_matches_reference(gffm):
    g = gffm
    A = dev(g, [[1, 2], [3, 4]], 65521, np.float32)
    with pytest.raises(g.ArgumentError):
        A * A
    A = dev(g, [[1, 2], [3, 4]], 2 ** 26 + 1, np.float64)
    with pytest.raises(g.ArgumentError):
        A * A


def test_argument_errors(gffm):
    g = gffm
    A = dev(g, np.ones((3, 4)), 7)
    B = dev(g, np.ones((3, 4)), 7)
    with pytest.raises(g.CuModArraySizeMismatchException):
        g.mul_(g.zeros(np.float64, 3, 4, 7), A, B)
    C = dev(g, np.ones((4, 2)), 11)
    with pytest.raises(g.CuModArrayModulusMismatchException):
        g.mul_(g.zeros(np.float64, 3, 2, 7), A, C)


def test_unreduced_inputs_are_reduced(gffm):
    g = gffm
    p = 101
    Ah = np.arange(-50, 50).reshape(10, 10) * 37
    Bh = np.arange(100).reshape(10, 10) * 1000
    A = g.CuModMatrix(Ah, p, mod=False, elem_type=np.float64)
    B = g.CuModMatrix(Bh, p, mod=False, elem_type=np.float64)
    for form in (1, 2, 3):
        C = g.zeros(np.float64, 10, 10, p)
        g.mul_(C, A, B, form=form)
        assert np.array_equal(g.Array(C).astype(np.int64), O.matmul_bigint(np.mod(Ah, p), np.mod(Bh, p), p))


@pytest.mark.parametrize("et", [np.float64, np.float32])
@pytest.mark.parametrize("p", [101, 251, 65521, 2 ** 26 - 5])
def test_unreduced_inputs_every_form(gffm, p, et):
    # waves that mix reduced elements with negative ones, ones >= p and large
    # magnitudes (exact in the storage type), in every form
    if et == np.float32 and p > 2897:
        pytest.skip("float32 storage: the reference's stripe rule admits p <= 2897")
    g = gffm
    rng = np.random.default_rng(p % 997)
    m, n, k = 130, 70, 300
    big = 2 ** 40 if et == np.float64 else 2 ** 23
    Ah = rng.integers(-big, big, size=(m, k))
    Bh = rng.integers(-3 * p, 3 * p, size=(k, n))
    Ah[::2, :] = np.mod(Ah[::2, :], p)
    Bh[:, ::3] = np.mod(Bh[:, ::3], p)
    Bh[5, :] = p
    A = g.CuModMatrix(Ah, p, mod=False, elem_type=et)
    B = g.CuModMatrix(Bh, p, mod=False, elem_type=et)
    ref = O.gemm(np.mod(Ah, p), np.mod(Bh, p), p)
    for form in (1, 2, 3):
        C = g.zeros(et, m, n, p)
        g.mul_(C, A, B, form=form)
        assert np.array_equal(g.Array(C).astype(np.int64), ref), f"form {form}"


def test_device_rand_matches_oracle_generator(gffm):
    g = gffm
    for p, et in ((65521, np.float64), (2 ** 26 - 5, np.float64), (127, np.float32)):
        X = g.rand(et, 300, 200, p, seed=5, offset=17)
        assert np.array_equal(g.Array(X).astype(np.int64), O.draw_matrix(300, 200, p, seed=5, offset=17))


@pytest.mark.parametrize("p", [2 ** 26 - 5, 127])
def test_config2_8192_vs_oracle(gffm, p):
    # full size: the 8192^3 product of the seeded residues (seed 1) in both
    # forms equals the oracle's product entry for entry (max |delta| = 0)
    g = gffm
    n = 8192
    print(f"config2 n={n} p={p}", flush=True)
    A = g.rand(np.float64, n, n, p, seed=1)
    B = g.rand(np.float64, n, n, p, seed=1, offset=n * n)
    Ah = g.Array(A).astype(np.int64)
    Bh = g.Array(B).astype(np.int64)
    ref = O.gemm(Ah, Bh, p)
    del Ah, Bh
    for form in (1, 2):
        C = g.zeros(np.float64, n, n, p)
        g.mul_(C, A, B, form=form)
        c = g.Array(C).astype(np.int64)
        del C
        assert int(np.abs(c - ref).max()) == 0, f"form {form}"
        del c


@pytest.mark.parametrize("p", [2, 127, 65521, 8388593, 2 ** 26 - 5])
@pytest.mark.parametrize("shape", [(1, 1), (100, 257), (257, 100), (5000, 3000), (64, 70000)], ids=str)
def test_gemv_matches_oracle(gffm, shape, p):
    # z = A x mod p through the streaming matrix-vector kernel (K slices
    # summed mod p for the long shapes); extreme residues in one row
    g = gffm
    m, k = shape
    print(f"gemv shape={shape} p={p}", flush=True)
    Ah = O.draw_matrix(m, k, p, seed=31)
    xh = O.draw_matrix(k, 1, p, seed=31, offset=m * k)
    Ah[0, :] = p - 1
    xh[: k // 2, 0] = p - 1
    z = g.zeros(np.float64, m, 1, p)
    g.mul_(z, dev(g, Ah, p), g.CuModVector(xh[:, 0], p, elem_type=np.float64))
    assert np.array_equal(g.Array(z).astype(np.int64), O.gemm(Ah, xh, p))


def test_gemv_float32_and_operator(gffm):
    g = gffm
    p = 2039
    Ah = O.draw_matrix(300, 400, p, seed=32)
    xh = O.draw_matrix(400, 1, p, seed=33)
    A = dev(g, Ah, p, np.float32)
    x = g.CuModVector(xh[:, 0], p, elem_type=np.float32)
    assert np.array_equal(g.Array(A * x).astype(np.int64), O.gemm(Ah, xh, p))
