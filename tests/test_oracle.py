"""Pins of the CPU oracle (no GPU): reference known answers, closed forms,
invariants, brute force on tiny inputs, and planted mistakes that the checks
must catch."""
import itertools
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))


@pytest.fixture(scope="module", autouse=True)
def _lib():
    O.lib()


# ---------------------------------------------------------------- GEMM

@pytest.mark.parametrize("case", GOLD["matmul"], ids=lambda c: c["name"])
def test_gemm_known_answers(case):
    C = O.gemm(case["A"], case["B"], case["p"])
    assert C.tolist() == case["C"]


@pytest.mark.parametrize("p", [2, 3, 127, 65521, 2 ** 23 - 15, 2 ** 26 - 5, 2 ** 26])
@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 5, 7), (17, 9, 33), (8, 8, 300)])
def test_gemm_vs_bigint(p, shape):
    m, n, k = shape
    A = O.draw_matrix(m, k, p, seed=11)
    B = O.draw_matrix(k, n, p, seed=11, offset=m * k)
    assert np.array_equal(O.gemm(A, B, p), O.matmul_bigint(A, B, p))


@pytest.mark.parametrize("p", [2 ** 26 - 5, 65521, 127])
def test_gemm_extreme_residues(p):
    # every entry p-1: the worst case for every overflow bound (Appendix: u64
    # periodic reduction); closed form (p-1)^2 * k mod p = k mod p
    k = 5000
    A = np.full((4, k), p - 1, dtype=np.int64)
    B = np.full((k, 3), p - 1, dtype=np.int64)
    C = O.gemm(A, B, p)
    assert np.all(C == (k % p))


def test_gemm_planted_transpose_is_caught():
    A = O.draw_matrix(5, 5, 101, seed=1)
    B = O.draw_matrix(5, 5, 101, seed=2)
    good = O.gemm(A, B, 101)
    wrong = O.gemm(B, A, 101)  # a transposed/operand-swapped product
    assert not np.array_equal(good, wrong)
    assert np.array_equal(good, O.matmul_bigint(A, B, 101))


# ---------------------------------------------------------------- generator

def test_generator_deterministic_and_in_range():
    X = O.draw_matrix(50, 40, 65521, seed=0)
    Y = O.draw_matrix(50, 40, 65521, seed=0)
    assert np.array_equal(X, Y)
    assert X.min() >= 0 and X.max() < 65521
    Z = O.draw_matrix(50, 40, 65521, seed=1)
    assert not np.array_equal(X, Z)
    # offset continues the same stream: A before B
    AB = O.draw_matrix(50, 80, 65521, seed=3)
    assert np.array_equal(AB[:, 40:], O.draw_matrix(50, 40, 65521, seed=3, offset=50 * 40))


def test_generator_uniformity_small_p():
    X = O.draw_matrix(200, 200, 7, seed=5).ravel()
    counts = np.bincount(X, minlength=7)
    assert counts.min() > 0.9 * X.size / 7 and counts.max() < 1.1 * X.size / 7


def test_splitmix_known_values():
    # SplitMix64 with state seed+gamma*(i+1): first outputs for seed 0 are the
    # published SplitMix64 test values; with p = 2^64 -> the raw output is
    # not representable, so check through p = 2^32 (high 32 bits of x*p)
    raw = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    for i, x in enumerate(raw):
        assert O.draw(0, i, 2 ** 32) == x >> 32


# ---------------------------------------------------------------- mod_inv / primes

@pytest.mark.parametrize("p", [2, 3, 7, 11, 251, 65521])
def test_mod_inv_exhaustive(p):
    for a in range(1, p):
        inv = O.mod_inv(a, p)
        assert 0 <= inv < p and (a * inv) % p == 1


def test_mod_inv_large():
    p = 2 ** 26 - 5
    for a in [1, 2, p - 1, 12345678, 2 ** 25]:
        assert (a * O.mod_inv(a, p)) % p == 1


def test_primes():
    small = [x for x in range(2, 2000) if all(x % d for d in range(2, int(x ** 0.5) + 1))]
    assert [x for x in range(2000) if O.is_prime(x)] == small
    assert O.is_prime(2 ** 26 - 5) and O.is_prime(65521) and not O.is_prime(2 ** 26 - 3)
    top = O.largest_primes_below(2 ** 23, 64)
    assert len(top) == 64 and top[0] == 8388593 and all(a > b for a, b in zip(top, top[1:]))


# ---------------------------------------------------------------- permutations

def test_permutation_semantics_reference_fixture():
    f = GOLD["permutation"]
    A = np.array(f["A"])
    st = [tuple(x) for x in f["stack"]]
    assert O.apply_col_perm(st, A).tolist() == f["col_perm"]
    assert O.apply_row_perm(st, A).tolist() == f["row_perm"]
    assert np.array_equal(O.apply_col_inv_perm(st, O.apply_col_perm(st, A)), A)
    assert np.array_equal(O.apply_row_inv_perm(st, O.apply_row_perm(st, A)), A)


def test_perm_stack_matrix_is_row_application():
    st = [(1, 3), (2, 3), (1, 2)]
    X = np.arange(9).reshape(3, 3)
    P = O.perm_stack_to_matrix(st, 3)
    assert np.array_equal(P @ X, O.apply_row_perm(st, X))
    # apply_col_inv_perm(S, X) = X * P_S (Q_S^{-1} = P_S)
    assert np.array_equal(X @ P, O.apply_col_inv_perm(st, X))


# ---------------------------------------------------------------- PLUQ

@pytest.mark.parametrize("case", GOLD["elimination"], ids=lambda c: c["name"])
def test_pluq_known_answers(case):
    A = np.mod(np.array(case["A"]), case["p"])
    for mode in (O.PLUQ_SRC, O.PLUQ_OPTC):
        r = O.pluq(A, case["p"], mode)
        assert r["rank"] == case["rank"]
        assert r["row_swaps"] == [tuple(x) for x in case["row_swaps"]]
        assert r["col_swaps"] == [tuple(x) for x in case["col_swaps"]]
        assert O.check_pluq(A, case["p"], r["U"], r["L"], r["row_swaps"], r["col_swaps"], r["rank"]) == []


def test_pluq_identity_is_trivial():
    I = np.eye(6, dtype=np.int64)
    r = O.pluq(I, 11, O.PLUQ_SRC)
    assert np.array_equal(r["U"], I) and np.array_equal(r["L"], I)


def _rank_deficient(m, n, r, p, seed):
    X = O.draw_matrix(m, r, p, seed)
    Y = O.draw_matrix(r, n, p, seed, offset=m * r)
    return O.gemm(X, Y, p)


SHAPES = [(1, 1), (1, 5), (5, 1), (6, 20), (20, 6), (10, 10), (33, 17), (17, 33), (40, 40)]


@pytest.mark.parametrize("p", [2, 3, 7, 251, 65521])
@pytest.mark.parametrize("shape", SHAPES)
def test_pluq_optc_always_valid(p, shape):
    m, n = shape
    for seed in range(4):
        A = O.draw_matrix(m, n, p, seed)
        if seed == 3 and min(m, n) > 2:
            A = _rank_deficient(m, n, min(m, n) // 2, p, seed)
        r = O.pluq(A, p, O.PLUQ_OPTC)
        assert O.check_pluq(A, p, r["U"], r["L"], r["row_swaps"], r["col_swaps"], r["rank"]) == []
        assert r["rank"] == O.rref(A, p)["rank"]


@pytest.mark.parametrize("p", [3, 7, 65521])
def test_src_equals_optc_when_src_valid(p):
    # the reference's pivot sequence: whenever src yields a valid factorisation
    # optC reproduces it bit for bit
    agree = 0
    for seed in range(30):
        m, n = 6 + seed % 5, 6 + (seed * 3) % 7
        A = O.draw_matrix(m, n, p, seed)
        if seed % 3 == 0:
            A = _rank_deficient(m, n, 3, p, seed)
        s = O.pluq(A, p, O.PLUQ_SRC)
        if s["L_extra"].any():
            continue
        if O.check_pluq(A, p, s["U"], s["L"], s["row_swaps"], s["col_swaps"], s["rank"]) != []:
            continue
        c = O.pluq(A, p, O.PLUQ_OPTC)
        for key in ("U", "L"):
            assert np.array_equal(s[key], c[key])
        assert s["row_swaps"] == c["row_swaps"] and s["col_swaps"] == c["col_swaps"] and s["rank"] == c["rank"]
        agree += 1
    assert agree >= 10


def test_src_defect_reproduced():
    # a zero column before the rank is exhausted: src skips the swapped-in
    # column and indexes L by column (pluq_kernels.jl:88,103,343), so its
    # output is not a factorisation; optC's is
    A = np.array([[0, 1, 1], [0, 1, 2], [0, 0, 1]], dtype=np.int64)
    s = O.pluq(A, 3, O.PLUQ_SRC)
    assert O.check_pluq(A, 3, s["U"], s["L"], s["row_swaps"], s["col_swaps"], s["rank"]) != []
    c = O.pluq(A, 3, O.PLUQ_OPTC)
    assert O.check_pluq(A, 3, c["U"], c["L"], c["row_swaps"], c["col_swaps"], c["rank"]) == []
    assert c["col_swaps"][0] == (1, 3)


def test_pluq_full_rank_is_partial_pivot_lu():
    # full-rank square: no column swaps; U unit upper, L lower with pivots on
    # the diagonal; the pivot is the column maximum (textbook partial pivoting
    # with the max-residue rule)
    p = 65521
    A = O.draw_matrix(30, 30, p, 9)
    r = O.pluq(A, p)
    assert r["col_swaps"] == [] and r["rank"] == 30
    # recompute the first pivot by hand
    first = int(np.argmax(A[:, 0]))
    if first != 0:
        assert r["row_swaps"][0] == (1, first + 1)
    assert r["L"][0, 0] == A[:, 0].max()


def test_pluq_gf2_brute_force():
    # every 3x3 matrix over GF(2): optC is a valid factorisation and its rank
    # equals the brute-force rank
    for bits in range(512):
        A = np.array([(bits >> i) & 1 for i in range(9)], dtype=np.int64).reshape(3, 3)
        r = O.pluq(A, 2, O.PLUQ_OPTC)
        assert O.check_pluq(A, 2, r["U"], r["L"], r["row_swaps"], r["col_swaps"], r["rank"]) == []
        rk = _gf2_rank(A)
        assert r["rank"] == rk


def _gf2_rank(A):
    rows = [int("".join(str(int(x)) for x in row), 2) for row in A]
    rank = 0
    for bit in reversed(range(A.shape[1])):
        piv = next((r for r in rows if (r >> bit) & 1), None)
        if piv is None:
            continue
        rows = [r ^ piv if (r >> bit) & 1 else r for r in rows if r != piv]
        rank += 1
    return rank


def test_check_pluq_catches_planted_mistakes():
    p = 251
    A = O.draw_matrix(8, 8, p, 4)
    r = O.pluq(A, p)
    L = r["L"].copy()
    L[[2, 5]] = L[[5, 2]]  # a swapped row
    assert O.check_pluq(A, p, r["U"], L, r["row_swaps"], r["col_swaps"], r["rank"]) != []
    U = r["U"].copy()
    U[1, 6] = (U[1, 6] + 1) % p  # an off-by-one entry
    assert O.check_pluq(A, p, U, r["L"], r["row_swaps"], r["col_swaps"], r["rank"]) != []
    bad_stack = list(r["row_swaps"])[::-1]  # stack in the wrong order
    if len(bad_stack) > 1 and bad_stack != r["row_swaps"]:
        assert O.check_pluq(A, p, r["U"], r["L"], bad_stack, r["col_swaps"], r["rank"]) != []


def test_check_pluq_freivalds_mode():
    p = 65521
    A = O.draw_matrix(60, 50, p, 2)
    r = O.pluq(A, p)
    assert O.check_pluq(A, p, r["U"], r["L"], r["row_swaps"], r["col_swaps"], r["rank"], exact=False) == []
    U = r["U"].copy()
    U[3, 40] = (U[3, 40] + 7) % p
    assert O.check_pluq(A, p, U, r["L"], r["row_swaps"], r["col_swaps"], r["rank"], exact=False) != []


# ---------------------------------------------------------------- LU (PLE)

@pytest.mark.parametrize("p", [2, 7, 65521])
@pytest.mark.parametrize("shape", SHAPES)
def test_lu_ple_invariants(p, shape):
    m, n = shape
    for seed in range(3):
        A = O.draw_matrix(m, n, p, seed)
        if seed == 2 and min(m, n) > 2:
            A = _rank_deficient(m, n, min(m, n) // 2, p, seed)
        r = O.lu(A, p)
        perm = r["perm"] - 1
        assert sorted(perm.tolist()) == list(range(m))
        assert np.array_equal(A[perm, :] % p, O.matmul_bigint(r["L"], r["U"], p))
        assert r["rank"] == O.rref(A, p)["rank"]
        # echelon: pivot columns strictly increasing, unit pivots
        pc = r["pivcols"] - 1
        assert all(a < b for a, b in zip(pc, pc[1:]))
        for i, c in enumerate(pc):
            assert r["U"][i, c] == 1 and not r["U"][i, :c].any()
        assert np.array_equal(pc + 1, O.rref(A, p)["pivcols"])


# ---------------------------------------------------------------- RREF

@pytest.mark.parametrize("p", [2, 3, 7, 251, 65521, 2 ** 26 - 5])
def test_rref_vs_python_gauss_jordan(p):
    for seed, (m, n) in enumerate([(1, 1), (4, 7), (7, 4), (9, 9), (12, 20), (20, 12)]):
        A = O.draw_matrix(m, n, p, seed)
        if seed % 2:
            A = _rank_deficient(m, n, max(1, min(m, n) // 3), p, seed)
        r = O.rref(A, p)
        R2, piv2 = O.rref_bigint(A, p)
        assert np.array_equal(r["R"], R2)
        assert r["pivcols"].tolist() == piv2
        ok, piv = O.is_rref(r["R"], p)
        assert ok and piv == piv2


def test_rref_row_space_and_rank_brute_force():
    # 2x3 over GF(3): compare against the row space computed by enumeration
    for seed in range(40):
        A = O.draw_matrix(2, 3, 3, seed)
        r = O.rref(A, 3)
        span_A = {tuple((a * A[0] + b * A[1]) % 3) for a in range(3) for b in range(3)}
        R = r["R"]
        span_R = {tuple((a * R[0] + b * R[1]) % 3) for a in range(3) for b in range(3)}
        assert span_A == span_R
        assert len(span_A) == 3 ** r["rank"]


def test_is_rref_catches_planted_mistake():
    A = O.draw_matrix(6, 9, 65521, 1)
    R = O.rref(A, 65521)["R"].copy()
    assert O.is_rref(R, 65521)[0]
    R[0, R.shape[1] - 1], R[1, R.shape[1] - 1] = R[1, R.shape[1] - 1], R[0, R.shape[1] - 1]
    R2 = R.copy()
    R2[1, 0] = 5  # a nonzero below a pivot
    assert not O.is_rref(R2, 65521)[0]


def _c5_instances():
    # the batched-elimination workload exactly as the benchmark and the GPU
    # tests build it
    from gpufinitefieldmatrices import workloads as W
    return W.elim_instances(4096, seed=4)


def test_c5_src_and_optc_differences():
    # how often the reference's pivot loop (src) and the build's Option C
    # semantics differ on the batched-elimination workload, and that every
    # difference is an instance where src is not a factorisation
    diff = invalid_src = 0
    for A, p in _c5_instances():
        s = O.pluq(A, p, O.PLUQ_SRC)
        c = O.pluq(A, p, O.PLUQ_OPTC)
        same = (np.array_equal(s["U"], c["U"]) and np.array_equal(s["L"], c["L"]) and not s["L_extra"].any()
                and s["row_swaps"] == c["row_swaps"] and s["col_swaps"] == c["col_swaps"]
                and s["rank"] == c["rank"])
        if same:
            continue
        diff += 1
        assert O.check_pluq(A, p, c["U"], c["L"], c["row_swaps"], c["col_swaps"], c["rank"]) == []
        if O.check_pluq(A, p, s["U"], s["L"], s["row_swaps"], s["col_swaps"], s["rank"]) != []:
            invalid_src += 1
    print(f"C5: src and Option C differ on {diff} of 4096 instances; src invalid on {invalid_src}")
    assert invalid_src == diff


@pytest.mark.parametrize("p", [2, 3, 7, 65521, 2 ** 26 - 5])
def test_oracle_inverse_is_an_inverse(p):
    # A * inv(A) = I by big-integer products; singular inputs give None
    for seed in range(6):
        n = 1 + seed * 7
        A = O.draw_matrix(n, n, p, seed)
        inv = O.inverse(A, p)
        rank = O.rref(A, p)["rank"]
        if rank < n:
            assert inv is None
            continue
        eye = np.eye(n, dtype=np.int64)
        assert (np.asarray(O.matmul_bigint(A, inv, p), dtype=np.int64) == eye).all()
        assert (np.asarray(O.matmul_bigint(inv, A, p), dtype=np.int64) == eye).all()
    S = O.draw_matrix(6, 6, p, 99)
    S[:, 5] = S[:, 0]
    assert O.inverse(S, p) is None


def test_oracle_inverse_reference_fixtures():
    # triangular_test.jl:26-31 and :74-79 (2x2 mod 7), and the identity
    U = np.array([[1, 3], [0, 2]])
    L = np.array([[1, 0], [3, 2]])
    for T in (U, L):
        inv = O.inverse(T, 7)
        assert ((T @ inv) % 7 == np.eye(2, dtype=np.int64)).all()
    assert (O.inverse(np.eye(50, dtype=np.int64), 2) == np.eye(50, dtype=np.int64)).all()
    # the upper inverse is upper triangular, the lower one lower triangular
    assert O.inverse(U, 7)[1, 0] == 0 and O.inverse(L, 7)[0, 1] == 0


# ---------------------------------------------------------------- pivot rule
# The reference picks the pivot with Base.findmax over the column
# (pluq_kernels.jl:188-189): the largest residue, the FIRST row among equal
# residues.  Two independent pins of the C restatement's choice: a pure-Python
# restatement whose search is numpy.argmax (documented first-occurrence rule),
# compared on every tie-heavy input, and an invariant checker that reads the
# rule off the factorisation itself.

def _same_pluq(a, b):
    return (a["rank"] == b["rank"] and a["row_swaps"] == b["row_swaps"] and a["col_swaps"] == b["col_swaps"]
            and np.array_equal(a["U"], b["U"]) and np.array_equal(a["L"], b["L"])
            and np.array_equal(a["L_extra"], b["L_extra"]))


def _same_lu(a, b):
    return (a["rank"] == b["rank"] and np.array_equal(a["perm"], b["perm"])
            and np.array_equal(a["pivcols"], b["pivcols"])
            and np.array_equal(a["U"], b["U"]) and np.array_equal(a["L"], b["L"]))


def _pivot_rule_errors(A, p, lib_=None):
    errs = []
    for mode in (O.PLUQ_OPTC, O.PLUQ_SRC):
        r = O.pluq(A, p, mode, lib_=lib_)
        if not _same_pluq(r, O.pluq_restated(A, p, mode)):
            errs.append(f"pluq mode {mode} differs from the restatement")
        if mode == O.PLUQ_OPTC:
            e = O.check_pivot_rule(r["L"], r["rank"], O.pivot_targets_from_stack(r["row_swaps"], r["rank"]))
            errs += [f"pluq: {x}" for x in e]
    r = O.lu(A, p, lib_=lib_)
    if not _same_lu(r, O.lu_restated(A, p)):
        errs.append("lu differs from the restatement")
    errs += [f"lu: {x}" for x in O.check_pivot_rule(r["L"], r["rank"], O.pivot_targets_from_perm(r["perm"], r["rank"]))]
    return errs


@pytest.mark.parametrize("p", [2, 3])
def test_pivot_rule_exhaustive_3x3(p):
    # every 3x3 matrix over GF(2) and GF(3): nearly every pivot is a tie
    for vals in itertools.product(range(p), repeat=9):
        A = np.array(vals, dtype=np.int64).reshape(3, 3)
        errs = _pivot_rule_errors(A, p)
        assert errs == [], (A.tolist(), errs)


@pytest.mark.parametrize("shape", [(2, 4), (4, 2), (3, 4)])
def test_pivot_rule_exhaustive_gf2_rectangular(shape):
    m, n = shape
    for vals in itertools.product(range(2), repeat=m * n):
        A = np.array(vals, dtype=np.int64).reshape(m, n)
        assert _pivot_rule_errors(A, 2) == [], A.tolist()


@pytest.mark.parametrize("p,shape", [(3, (8, 8)), (7, (8, 8)), (3, (6, 9)), (7, (9, 6)), (2, (10, 10))])
def test_pivot_rule_random_tie_heavy(p, shape):
    rng = np.random.default_rng(p * 100 + shape[0])
    for t in range(300):
        A = rng.integers(0, p, size=shape)
        if t % 3 == 0:
            A[:, rng.integers(0, shape[1])] = 0          # a zero column
        if t % 5 == 0:
            A[rng.integers(0, shape[0]), :] = A[rng.integers(0, shape[0]), :]  # a repeated row
        assert _pivot_rule_errors(A, p) == [], (t, A.tolist())


def test_pivot_rule_check_medium_sizes():
    # the invariant checker on the C oracle at sizes the Python restatement
    # cannot reach; it is what the GPU tests apply at 4096^2
    for p, m, n in ((3, 200, 150), (65521, 300, 300), (2, 128, 200)):
        A = O.draw_matrix(m, n, p, seed=p + m)
        r = O.pluq(A, p)
        assert O.check_pivot_rule(r["L"], r["rank"], O.pivot_targets_from_stack(r["row_swaps"], r["rank"])) == []
        l = O.lu(A, p)
        assert O.check_pivot_rule(l["L"], l["rank"], O.pivot_targets_from_perm(l["perm"], l["rank"])) == []


@pytest.fixture(scope="module")
def planted_tie_last(tmp_path_factory):
    # the oracle rebuilt with the tie broken to the LAST row (a `>=` in its
    # column argmax); the checks above must reject it
    import subprocess
    src = os.path.join(os.path.dirname(HERE), "oracle", "modp_oracle.c")
    out = str(tmp_path_factory.mktemp("planted") / "liboracle_tie_last.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-std=gnu11", "-shared", "-DORACLE_PLANT_TIE_LAST",
                    "-o", out, src], check=True)
    return O.load_variant(out)


def test_planted_tie_break_is_caught(planted_tie_last):
    A = np.array([[1, 0, 1], [1, 1, 0], [0, 1, 1]], dtype=np.int64)
    errs = _pivot_rule_errors(A, 2, lib_=planted_tie_last)
    assert any("differs" in e for e in errs) and any("tie" in e for e in errs), errs
    # and on the medium-size check, where only the invariant is available
    B = O.draw_matrix(200, 150, 3, seed=5)
    r = O.pluq(B, 3, lib_=planted_tie_last)
    assert O.check_pivot_rule(r["L"], r["rank"], O.pivot_targets_from_stack(r["row_swaps"], r["rank"])) != []
    l = O.lu(B, 3, lib_=planted_tie_last)
    assert O.check_pivot_rule(l["L"], l["rank"], O.pivot_targets_from_perm(l["perm"], l["rank"])) != []


def test_check_pivot_rule_catches_a_smaller_pivot():
    # a valid factorisation with a non-maximal pivot: P = I on [[1,0],[2,1]]
    # mod 5 (pivot 1 although 2 is larger)
    L = np.array([[1, 0], [2, 1]], dtype=np.int64)
    assert O.check_pivot_rule(L, 2, [0, 1]) != []
