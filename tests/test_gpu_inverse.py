"""Triangular and matrix inverses on the GPU through the C ABI: equal to the
oracle's Gauss-Jordan inverse (inverses are unique), A * A^{-1} = I, the
reference's triangular fixtures, and the error cases."""
import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))


def dev(g, X, p, et=np.float64):
    return g.CuModMatrix(np.asarray(X), p, elem_type=et)


def rand_tri(n, p, seed, upper, cols=None):
    rng = np.random.default_rng(seed)
    cols = n if cols is None else cols
    T = rng.integers(0, p, size=(n, cols))
    T = np.triu(T) if upper else np.tril(T)
    for i in range(min(n, cols)):
        T[i, i] = rng.integers(1, p)
    return T


@pytest.mark.parametrize("p", [2, 7, 65521, 2 ** 26 - 5])
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 130, 257])
@pytest.mark.parametrize("upper", [True, False], ids=["upper", "lower"])
def test_triangular_inverse_matches_oracle(gffm, upper, n, p):
    g = gffm
    print(f"trtri upper={upper} n={n} p={p}", flush=True)
    T = rand_tri(n, p, seed=n * 7 + (1 if upper else 0), upper=upper)
    f = g.upper_triangular_inverse_no_copy if upper else g.lower_triangular_inverse_no_copy
    X = g.Array(f(dev(g, T, p))).astype(np.int64)
    assert np.array_equal(X, O.inverse(T, p))


def test_triangular_ignores_the_other_triangle(gffm):
    # entries outside the triangle are not read
    g = gffm
    p = 251
    T = rand_tri(100, p, seed=3, upper=True)
    junk = T + np.tril(np.random.default_rng(4).integers(0, p, size=(100, 100)), -1)
    X = g.Array(g.upper_triangular_inverse_no_copy(dev(g, junk, p))).astype(np.int64)
    assert np.array_equal(X, O.inverse(T, p))


def test_reference_triangular_fixtures(gffm):
    # triangular_test.jl: 2x2 mod 7 upper and lower, identities mod 2
    g = gffm
    for T, f in ((np.array([[1, 3], [0, 2]]), g.upper_triangular_inverse_no_copy),
                 (np.array([[1, 0], [3, 2]]), g.lower_triangular_inverse_no_copy)):
        A = dev(g, T, 7)
        assert np.array_equal(g.Array(A * f(A)).astype(np.int64), np.eye(2, dtype=np.int64))
    E = dev(g, np.eye(300, dtype=np.int64), 2)
    for f in (g.upper_triangular_inverse_no_copy, g.lower_triangular_inverse_no_copy):
        assert np.array_equal(g.Array(f(E)).astype(np.int64), np.eye(300, dtype=np.int64))


def test_wide_upper_right_inverse(gffm):
    # triangular_test.jl:42-45 shape (1000 x 5000 mod 13 there): A * A_inv = I
    g = gffm
    p = 13
    T = np.triu(np.random.default_rng(5).integers(1, 13, size=(200, 700)))
    A = dev(g, T, p)
    X = g.upper_triangular_inverse_no_copy(A)
    assert X.size == (700, 200)
    assert np.array_equal(g.Array(A * X).astype(np.int64), np.eye(200, dtype=np.int64))
    Xh = g.Array(X).astype(np.int64)
    assert not Xh[200:, :].any()
    assert np.array_equal(Xh[:200, :], O.inverse(T[:, :200], p))


def test_tall_lower_raises(gffm):
    g = gffm
    with pytest.raises(g.InverseNotDefinedException):
        g.lower_triangular_inverse_no_copy(dev(g, np.tril(np.ones((5, 3), dtype=np.int64)), 7))


def test_zero_diagonal_raises(gffm):
    g = gffm
    T = rand_tri(70, 7, seed=9, upper=True)
    T[66, 66] = 0
    with pytest.raises(g.MatrixNotInvertibleException):
        g.upper_triangular_inverse_no_copy(dev(g, T, 7))


@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("p", [3, 7, 251, 65521, 2 ** 26 - 5])
@pytest.mark.parametrize("n", [1, 5, 64, 65, 200, 513])
def test_inverse_matches_oracle(gffm, n, p, form):
    g = gffm
    print(f"inverse n={n} p={p} form={form}", flush=True)
    A = O.draw_matrix(n, n, p, seed=n + p % 1000)
    ref = O.inverse(A, p)
    ok, X = g.is_invertible_with_inverse(dev(g, A, p), form=form)
    assert ok == (ref is not None)
    assert g.is_invertible(dev(g, A, p)) == ok
    if ok:
        assert np.array_equal(g.Array(X).astype(np.int64), ref)


def test_inverse_de_rham_fixture(gffm):
    # de_rham_test.jl: the 10x10 matrix mod 7 is invertible
    g = gffm
    case = [c for c in GOLD["elimination"] if "de_rham" in c["src"]][0]
    A = np.array(case["A"], dtype=np.int64)
    p = int(case["p"])
    X = g.inverse(dev(g, A, p))
    assert np.array_equal(g.Array(dev(g, A, p) * X).astype(np.int64), np.eye(A.shape[0], dtype=np.int64))
    assert np.array_equal(g.Array(X).astype(np.int64), O.inverse(A, p))


def test_singular_inputs(gffm):
    g = gffm
    p = 65521
    A = O.draw_matrix(120, 120, p, seed=12)
    A[:, 7] = (A[:, 3] * 5 + A[:, 9]) % p
    M = dev(g, A, p)
    assert not g.is_invertible(M)
    assert g.is_invertible_with_inverse(M) == (False, None)
    with pytest.raises(g.MatrixNotInvertibleException):
        g.inverse(M)
    with pytest.raises(g.CuModMatrixNotSquareException):
        g.inverse(dev(g, O.draw_matrix(4, 5, p, seed=1), p))


def test_inverse_float32_storage(gffm):
    g = gffm
    p = 251
    A = O.draw_matrix(100, 100, p, seed=13)
    X = g.inverse(dev(g, A, p, np.float32))
    assert np.array_equal(g.Array(X).astype(np.int64), O.inverse(A, p))


def test_inverse_4096_freivalds(gffm):
    # full size: A * (A^{-1} x) = x for random x, with exact host arithmetic
    g = gffm
    p, n = 65521, 4096
    A = g.rand(np.float64, n, n, p, seed=2)
    X = g.inverse(A)
    Ah = g.Array(A).astype(np.int64)
    Xh = g.Array(X).astype(np.int64)
    rng = np.random.default_rng(0)
    for _ in range(2):
        x = rng.integers(0, p, size=n)
        y = O._mod_matvec(Xh, x, p)
        assert np.array_equal(O._mod_matvec(Ah, y, p), x)


def test_inverse_4096_small_prime_freivalds(gffm):
    # p <= 255 at full size: PLUQ with 8-bit residues and the triangular
    # inverses' products in the int8 digit form
    g = gffm
    p, n = 251, 4096
    for seed in (2, 3, 4):
        A = g.rand(np.float64, n, n, p, seed=seed)
        if g.is_invertible(A):
            break
    X = g.inverse(A)
    Ah = g.Array(A).astype(np.int64)
    Xh = g.Array(X).astype(np.int64)
    rng = np.random.default_rng(1)
    for _ in range(2):
        x = rng.integers(0, p, size=n)
        y = O._mod_matvec(Xh, x, p)
        assert np.array_equal(O._mod_matvec(Ah, y, p), x)


@pytest.mark.parametrize("p1,p2", [(2, 3), (65521, 65519), (67108859, 67108837), (8388593, 7)])
def test_crt_two_moduli(gffm, p1, p2):
    # per-prime products combined: CRT(A*B mod p1, A*B mod p2) = A*B mod p1*p2
    # for small integer entries, against Python integers
    g = gffm
    rng = np.random.default_rng(p1 % 1000)
    A = rng.integers(0, 50, size=(40, 30))
    B = rng.integers(0, 50, size=(30, 20))
    P1 = g.Array(dev(g, A, p1) * dev(g, B, p1)).astype(np.int64)
    P2 = g.Array(dev(g, A, p2) * dev(g, B, p2)).astype(np.int64)
    Y = g.crt(dev(g, P1, p1), dev(g, P2, p2))
    assert Y.N == p1 * p2
    exact = (A.astype(object) @ B.astype(object)) % (p1 * p2)
    assert np.array_equal(g.Array(Y).astype(np.int64), np.asarray(exact, dtype=np.int64))
    # random residue pairs: both congruences hold
    X1 = rng.integers(0, p1, size=(17, 9))
    X2 = rng.integers(0, p2, size=(17, 9))
    Yh = g.Array(g.crt(dev(g, X1, p1), dev(g, X2, p2))).astype(np.int64)
    assert (Yh % p1 == X1).all() and (Yh % p2 == X2).all() and (Yh < p1 * p2).all()


def test_crt_errors(gffm):
    g = gffm
    X = dev(g, np.ones((3, 3), dtype=np.int64), 6)
    with pytest.raises(g.ArgumentError):
        g.crt(X, dev(g, np.ones((3, 3), dtype=np.int64), 4))     # not coprime
    with pytest.raises(g.CuModArraySizeMismatchException):
        g.crt(X, dev(g, np.ones((3, 4), dtype=np.int64), 5))
