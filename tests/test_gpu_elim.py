"""Elimination parity on the GPU: pluq / lu / rref bit-exact against the
oracle (same U, L, stacks, rank, RREF), invariants at full size."""
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


def rank_deficient(m, n, r, p, seed):
    X = O.draw_matrix(m, r, p, seed)
    Y = O.draw_matrix(r, n, p, seed, offset=m * r)
    return O.gemm(X, Y, p)


def check_pluq_equal(g, Ah, p, form=0, et=np.float64):
    A = dev(g, Ah, p, et)
    U, L, P, Q, rank = g.pluq(A, form=form)
    ref = O.pluq(Ah, p, O.PLUQ_OPTC)
    assert rank == ref["rank"]
    assert P == ref["row_swaps"]
    assert Q == ref["col_swaps"]
    assert np.array_equal(g.Array(U).astype(np.int64), ref["U"])
    assert np.array_equal(g.Array(L).astype(np.int64), ref["L"])
    # the input is not modified
    assert np.array_equal(g.Array(A).astype(np.int64), np.mod(Ah, p))


@pytest.mark.parametrize("case", GOLD["elimination"], ids=lambda c: c["name"])
def test_pluq_known_answers(gffm, case):
    g = gffm
    p = case["p"]
    Ah = np.mod(np.array(case["A"]), p)
    U, L, P, Q = g.pluq_gpu_kernel(dev(g, Ah, p))
    assert P == [tuple(x) for x in case["row_swaps"]]
    assert Q == [tuple(x) for x in case["col_swaps"]]
    assert O.check_pluq(Ah, p, g.Array(U).astype(np.int64), g.Array(L).astype(np.int64), P, Q, case["rank"]) == []


SHAPES = [(1, 1), (1, 7), (7, 1), (6, 20), (20, 6), (10, 10), (33, 17), (17, 33), (64, 64), (100, 37),
          (37, 100), (129, 130), (200, 200)]


@pytest.mark.parametrize("p", [2, 3, 7, 251, 65521, 2 ** 26 - 5])
@pytest.mark.parametrize("shape", SHAPES, ids=str)
def test_pluq_vs_oracle(gffm, shape, p):
    m, n = shape
    for seed in range(3):
        print(f"pluq shape={shape} p={p} seed={seed}", flush=True)
        Ah = O.draw_matrix(m, n, p, seed)
        if seed == 2 and min(m, n) > 2:
            Ah = rank_deficient(m, n, min(m, n) // 2, p, seed)
        check_pluq_equal(gffm, Ah, p)


def test_pluq_zero_columns_mid_matrix(gffm):
    # zero columns before the rank is exhausted exercise the column swaps
    p = 7
    for seed in range(6):
        Ah = O.draw_matrix(40, 50, p, seed)
        Ah[:, [0, 3, 4, 17, 33]] = 0
        Ah[:, 45] = 0
        print(f"pluq zero columns seed={seed}", flush=True)
        check_pluq_equal(gffm, Ah, p)


def test_pluq_all_zero_and_single_nonzero(gffm):
    check_pluq_equal(gffm, np.zeros((5, 9), dtype=np.int64), 11)
    Z = np.zeros((9, 5), dtype=np.int64)
    Z[7, 3] = 4
    check_pluq_equal(gffm, Z, 11)


@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("n,r", [(512, 512), (700, 300), (1024, 1024), (1500, 1100)])
def test_pluq_medium_vs_oracle(gffm, n, r, form):
    p = 65521
    print(f"pluq medium n={n} r={r} form={form}", flush=True)
    Ah = O.draw_matrix(n, n, p, seed=2) if r == n else rank_deficient(n, n, r, p, seed=2)
    check_pluq_equal(gffm, Ah, p, form=form)


@pytest.mark.parametrize("shape,r", [((4200, 200), 200), ((4300, 150), 90), ((9000, 140), 140)], ids=str)
def test_pluq_lu_tall_vs_oracle(gffm, shape, r):
    # more than 4096 rows: two-column panels, 64 panels per outer block
    g = gffm
    p = 65521
    m, n = shape
    print(f"tall pluq/lu shape={shape} r={r}", flush=True)
    Ah = O.draw_matrix(m, n, p, seed=9) if r == n else rank_deficient(m, n, r, p, seed=9)
    check_pluq_equal(g, Ah, p, form=1)
    U, L, perm, piv, rank = g.lu_full(dev(g, Ah, p))
    ref = O.lu(Ah, p)
    assert rank == ref["rank"] == r
    assert perm == ref["perm"].tolist() and piv == ref["pivcols"].tolist()
    assert np.array_equal(g.Array(U).astype(np.int64), ref["U"])
    assert np.array_equal(g.Array(L).astype(np.int64), ref["L"])


@pytest.mark.parametrize("p", [8388593, 2 ** 26 - 5])
@pytest.mark.parametrize("kind", ["full", "deficient", "zero_cols"])
def test_medium_large_prime_all_modes(gffm, p, kind):
    # panels, block updates and block ends on the reduce-every-step path
    # (p >= 2^16), several outer blocks, every elimination mode
    g = gffm
    m, n = 600, 520
    print(f"large-p medium kind={kind} p={p}", flush=True)
    if kind == "full":
        Ah = O.draw_matrix(m, n, p, seed=21)
    elif kind == "deficient":
        Ah = rank_deficient(m, n, 300, p, seed=21)
    else:
        Ah = O.draw_matrix(m, n, p, seed=21)
        Ah[:, [5, 130, 131, 300]] = 0
    check_pluq_equal(g, Ah, p, form=0)
    U, L, perm, piv, rank = g.lu_full(dev(g, Ah, p))
    ref = O.lu(Ah, p)
    assert rank == ref["rank"] and perm == ref["perm"].tolist() and piv == ref["pivcols"].tolist()
    assert np.array_equal(g.Array(U).astype(np.int64), ref["U"])
    assert np.array_equal(g.Array(L).astype(np.int64), ref["L"])
    R, piv, rank = g.rref_full(dev(g, Ah, p))
    ref = O.rref(Ah, p)
    assert rank == ref["rank"] and piv == ref["pivcols"].tolist()
    assert np.array_equal(g.Array(R).astype(np.int64), ref["R"])


def test_pluq_float32_storage(gffm):
    for p in (7, 251, 65521):
        Ah = O.draw_matrix(60, 45, p, seed=4)
        check_pluq_equal(gffm, Ah, p, et=np.float32)


@pytest.mark.parametrize("p", [2, 7, 65521])
@pytest.mark.parametrize("shape", [(1, 1), (6, 20), (20, 6), (33, 33), (100, 37), (37, 100), (150, 150)], ids=str)
def test_lu_vs_oracle(gffm, shape, p):
    g = gffm
    m, n = shape
    for seed in range(3):
        Ah = O.draw_matrix(m, n, p, seed)
        if seed == 2 and min(m, n) > 2:
            Ah = rank_deficient(m, n, min(m, n) // 2, p, seed)
        if seed == 1 and n > 5:
            Ah[:, [1, 4]] = 0
        print(f"lu shape={shape} p={p} seed={seed}", flush=True)
        U, L, perm, piv, rank = g.lu_full(dev(g, Ah, p))
        ref = O.lu(Ah, p)
        assert rank == ref["rank"]
        assert perm == ref["perm"].tolist()
        assert piv == ref["pivcols"].tolist()
        assert np.array_equal(g.Array(U).astype(np.int64), ref["U"])
        assert np.array_equal(g.Array(L).astype(np.int64), ref["L"])


@pytest.mark.parametrize("p", [2, 3, 251, 65521, 2 ** 26 - 5])
@pytest.mark.parametrize("shape", [(1, 1), (5, 9), (9, 5), (64, 64), (70, 130), (130, 70), (300, 300)], ids=str)
def test_rref_vs_oracle(gffm, shape, p):
    g = gffm
    m, n = shape
    for seed in range(2):
        Ah = O.draw_matrix(m, n, p, seed)
        if seed == 1 and min(m, n) > 2:
            Ah = rank_deficient(m, n, min(m, n) // 3, p, seed)
        print(f"rref shape={shape} p={p} seed={seed}", flush=True)
        R, piv, rank = g.rref_full(dev(g, Ah, p))
        ref = O.rref(Ah, p)
        assert rank == ref["rank"] and piv == ref["pivcols"].tolist()
        assert np.array_equal(g.Array(R).astype(np.int64), ref["R"])


def test_rref_medium_rank_deficient(gffm):
    g = gffm
    p = 65521
    Ah = rank_deficient(800, 900, 500, p, seed=6)
    R, piv, rank = g.rref_full(dev(g, Ah, p))
    ref = O.rref(Ah, p)
    assert rank == 500 and np.array_equal(g.Array(R).astype(np.int64), ref["R"])


def test_not_prime_modulus_is_an_error(gffm):
    g = gffm
    A = dev(g, [[2, 1], [1, 1]], 4)
    with pytest.raises(g.CuModMatrixModulusNotPrimeException):
        g.pluq(A)
    # mid-panel: 20 unit pivots, then only even residues mod 4 (pivot 2 has
    # no inverse) -- every mode raises, lu and rref from their device-cursor
    # panels
    rng = np.random.default_rng(8)
    Ah = np.zeros((200, 150), dtype=np.int64)
    Ah[:20, :20] = np.eye(20, dtype=np.int64)
    Ah[:, 20:] = 2 * rng.integers(0, 2, size=(200, 130))
    Ah[25, 20] = 2
    for f in (g.pluq, g.lu_full, g.rref_full):
        with pytest.raises(g.CuModMatrixModulusNotPrimeException):
            f(dev(g, Ah, 4))


_C3 = {}


def _config3_input(g, kind):
    """The 4096^2 elimination inputs mod 65521: uniform residues (seed 2,
    full rank) or the product of a 4096x3000 and a 3000x4096 uniform factor
    (seed 2, rank 3000), with the product checked against the oracle."""
    p, n = 65521, 4096
    if kind not in _C3:
        if kind == "full":
            A = g.rand(np.float64, n, n, p, seed=2)
        else:
            X = g.rand(np.float64, n, 3000, p, seed=2)
            Y = g.rand(np.float64, 3000, n, p, seed=2, offset=n * 3000)
            A = X * Y
            assert np.array_equal(g.Array(A).astype(np.int64),
                                  O.gemm(g.Array(X).astype(np.int64), g.Array(Y).astype(np.int64), p))
        _C3[kind] = (A, g.Array(A).astype(np.int64))
    return _C3[kind]


@pytest.mark.parametrize("kind", ["full", "rank3000"])
@pytest.mark.parametrize("form", [1, 2])
def test_config3_4096_pluq_vs_oracle(gffm, kind, form):
    # full size: U, L, both swap stacks and the rank equal the oracle's, and
    # the pivots follow the reference's rule (largest residue, first row)
    g = gffm
    p = 65521
    print(f"config3 pluq {kind} form {form}", flush=True)
    A, Ah = _config3_input(g, kind)
    U, L, P, Q, rank = g.pluq(A, form=form)
    ref = O.pluq(Ah, p, O.PLUQ_OPTC)
    assert rank == ref["rank"] == (4096 if kind == "full" else 3000)
    assert P == ref["row_swaps"] and Q == ref["col_swaps"]
    Lh = g.Array(L).astype(np.int64)
    assert np.array_equal(g.Array(U).astype(np.int64), ref["U"])
    assert np.array_equal(Lh, ref["L"])
    assert O.check_pivot_rule(Lh, rank, O.pivot_targets_from_stack(P, rank)) == []


@pytest.mark.parametrize("kind", ["full", "rank3000"])
def test_config3_4096_lu_vs_oracle(gffm, kind):
    g = gffm
    p = 65521
    print(f"config3 lu {kind}", flush=True)
    A, Ah = _config3_input(g, kind)
    U, L, perm, piv, rank = g.lu_full(A)
    ref = O.lu(Ah, p)
    assert rank == ref["rank"]
    assert list(perm) == ref["perm"].tolist() and list(piv) == ref["pivcols"].tolist()
    Lh = g.Array(L).astype(np.int64)
    assert np.array_equal(g.Array(U).astype(np.int64), ref["U"])
    assert np.array_equal(Lh, ref["L"])
    assert O.check_pivot_rule(Lh, rank, O.pivot_targets_from_perm(perm, rank)) == []


@pytest.mark.parametrize("kind", ["full", "rank3000"])
def test_config3_4096_rref_vs_oracle(gffm, kind):
    g = gffm
    p = 65521
    print(f"config3 rref {kind}", flush=True)
    A, Ah = _config3_input(g, kind)
    R, piv, rank = g.rref_full(A)
    ref = O.rref(Ah, p)
    assert rank == ref["rank"] and list(piv) == ref["pivcols"].tolist()
    assert np.array_equal(g.Array(R).astype(np.int64), ref["R"])
