"""Eliminations through the W-resident panel kernel: every panel forced onto
it (MODP_PANEL_GLOBAL=1) at sizes the register panel also handles, and
matrices taller than the register panel's 16,384 rows, bit-exact against the
oracle (same U, L, stacks, rank, RREF)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def dev(g, X, p, et=np.float64):
    return g.CuModMatrix(np.asarray(X), p, elem_type=et)


def rank_deficient(m, n, r, p, seed):
    X = O.draw_matrix(m, r, p, seed)
    Y = O.draw_matrix(r, n, p, seed, offset=m * r)
    return O.gemm(X, Y, p)


def all_modes_equal(g, Ah, p):
    A = dev(g, Ah, p)
    U, L, P, Q, rank = g.pluq(A, form=0)
    ref = O.pluq(Ah, p, O.PLUQ_OPTC)
    assert rank == ref["rank"]
    assert P == ref["row_swaps"] and Q == ref["col_swaps"]
    assert np.array_equal(g.Array(U).astype(np.int64), ref["U"])
    assert np.array_equal(g.Array(L).astype(np.int64), ref["L"])
    U, L, perm, piv, rank = g.lu_full(dev(g, Ah, p))
    ref = O.lu(Ah, p)
    assert rank == ref["rank"] and perm == ref["perm"].tolist() and piv == ref["pivcols"].tolist()
    assert np.array_equal(g.Array(U).astype(np.int64), ref["U"])
    assert np.array_equal(g.Array(L).astype(np.int64), ref["L"])
    R, piv, rank = g.rref_full(dev(g, Ah, p))
    ref = O.rref(Ah, p)
    assert rank == ref["rank"] and piv == ref["pivcols"].tolist()
    assert np.array_equal(g.Array(R).astype(np.int64), ref["R"])


CASES = [((5, 9), "full"), ((9, 5), "full"), ((70, 40), "zero_cols"), ((300, 250), "full"),
         ((400, 380), "deficient"), ((260, 300), "zero_cols"), ((4300, 150), "deficient")]


@pytest.mark.parametrize("p", [7, 65521, 2 ** 26 - 5])
@pytest.mark.parametrize("shape,kind", CASES, ids=str)
def test_forced_global_panel_vs_oracle(gffm, monkeypatch, shape, kind, p):
    monkeypatch.setenv("MODP_PANEL_GLOBAL", "1")
    m, n = shape
    print(f"global panel shape={shape} kind={kind} p={p}", flush=True)
    if kind == "full":
        Ah = O.draw_matrix(m, n, p, seed=31)
    elif kind == "deficient":
        Ah = rank_deficient(m, n, min(m, n) // 2, p, seed=31)
    else:
        Ah = O.draw_matrix(m, n, p, seed=31)
        Ah[:, [0, 3, 17, min(n - 1, 33)]] = 0
    all_modes_equal(gffm, Ah, p)


@pytest.mark.parametrize("shape,r,p", [((17000, 64), 64, 65521), ((20000, 48), 30, 2 ** 26 - 5),
                                       ((16500, 40), 40, 7)], ids=str)
def test_more_than_16384_rows_vs_oracle(gffm, shape, r, p):
    m, n = shape
    print(f"tall elimination shape={shape} r={r} p={p}", flush=True)
    Ah = O.draw_matrix(m, n, p, seed=41) if r == n else rank_deficient(m, n, r, p, seed=41)
    all_modes_equal(gffm, Ah, p)
