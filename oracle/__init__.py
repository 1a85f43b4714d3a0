"""CPU oracle for mod-p GEMM and mod-p elimination -- TEST INFRASTRUCTURE ONLY.

Only the tests, the package smoke check and the benchmark's CPU-baseline leg
import this module, and only as the checker.  The product (the HIP library and
its Python host mirror) never imports, links or falls back to it.

Two layers:

* ctypes bindings to ``liboracle.so`` (built from ``modp_oracle.c`` by the
  Makefile next to this file): the C restatement of the reference
  arithmetic (see that file's header for the file:line each function follows);
* independent pure-Python checks used to pin the C restatement: big-integer
  matrix products, an invariant checker for PLUQ / PLE factorisations, an
  RREF-form checker, Freivalds-style randomised product checks, and the
  transposition-stack semantics of the reference's permutations.jl.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PLUQ_SRC = 0
PLUQ_OPTC = 1

_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_p64 = ctypes.POINTER(ctypes.c_int64)


def _bind(L):
    """Declare the C signatures on a loaded liboracle (or a test-built variant)."""
    L.oracle_draw.restype = _u64
    L.oracle_draw.argtypes = [_u64, _u64, _u64]
    L.oracle_fill.restype = None
    L.oracle_fill.argtypes = [_p64, _i64, _i64, _i64, _u64, _u64, _u64]
    L.oracle_mod_inv.restype = _i64
    L.oracle_mod_inv.argtypes = [_i64, _i64]
    L.oracle_gemm.restype = None
    L.oracle_gemm.argtypes = [_i64, _i64, _i64, _p64, _i64, _p64, _i64, _p64, _i64, _i64]
    L.oracle_pluq.restype = ctypes.c_int
    L.oracle_pluq.argtypes = [ctypes.c_int, _i64, _i64, _p64, _i64, _i64,
                              _p64, _i64, _p64, _i64, _p64, _p64, _p64, _p64, _i64, _p64]
    L.oracle_lu.restype = ctypes.c_int
    L.oracle_lu.argtypes = [_i64, _i64, _p64, _i64, _i64, _p64, _i64, _p64, _i64,
                            _p64, _p64, _p64]
    L.oracle_rref.restype = ctypes.c_int
    L.oracle_rref.argtypes = [_i64, _i64, _p64, _i64, _i64, _p64, _i64, _p64, _p64]
    L.oracle_is_prime.restype = ctypes.c_int
    L.oracle_is_prime.argtypes = [_u64]
    return L


def lib():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built: run `make -C oracle`")
        _LIB = _bind(ctypes.CDLL(path))
    return _LIB


def load_variant(path: str):
    """Load another build of modp_oracle.c (the tests build planted-mistake
    variants with -D switches) and return its bound library, to pass as
    ``lib_`` to pluq / lu."""
    return _bind(ctypes.CDLL(path))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_p64)


def _as_f64(a) -> np.ndarray:
    """int64 column-major copy of a 2-D array-like."""
    return np.asfortranarray(np.asarray(a, dtype=np.int64))


# ---------------------------------------------------------------------------
# generator
# ---------------------------------------------------------------------------

def draw(seed: int, index: int, p: int) -> int:
    return int(lib().oracle_draw(seed, index, p))


def draw_matrix(rows: int, cols: int, p: int, seed: int, offset: int = 0) -> np.ndarray:
    """rows x cols matrix of draws, column-major element index + offset."""
    X = np.zeros((rows, cols), dtype=np.int64, order="F")
    if rows and cols:
        lib().oracle_fill(_ptr(X), rows, cols, rows, seed, offset, p)
    return X


# ---------------------------------------------------------------------------
# arithmetic
# ---------------------------------------------------------------------------

def mod_inv(a: int, p: int) -> int:
    return int(lib().oracle_mod_inv(a, p))


def gemm(A, B, p: int) -> np.ndarray:
    A = _as_f64(A); B = _as_f64(B)
    m, k = A.shape
    k2, n = B.shape
    assert k == k2
    C = np.zeros((m, n), dtype=np.int64, order="F")
    if m and n:
        lib().oracle_gemm(m, n, k, _ptr(A), max(m, 1), _ptr(B), max(k, 1), _ptr(C), max(m, 1), p)
    return C


def pluq(A, p: int, mode: int = PLUQ_OPTC, lib_=None) -> dict:
    """PLUQ of A mod p.  Returns dict(U, L, row_swaps, col_swaps, rank).

    L is m x m (for src mode the columns beyond m that the reference's
    col-indexed L can touch are returned separately as L_extra)."""
    A = _as_f64(A)
    m, n = A.shape
    lcols = max(m, n)
    cap = 2 * n + m + 4
    U = np.zeros((m, n), dtype=np.int64, order="F")
    L = np.zeros((m, lcols), dtype=np.int64, order="F")
    rs = np.zeros(2 * cap, dtype=np.int64)
    cs = np.zeros(2 * cap, dtype=np.int64)
    nr = _i64(); nc = _i64(); rank = _i64()
    rc = (lib_ or lib()).oracle_pluq(mode, m, n, _ptr(A), max(m, 1), p, _ptr(U), max(m, 1), _ptr(L), max(m, 1),
                           _ptr(rs), ctypes.byref(nr), _ptr(cs), ctypes.byref(nc), cap, ctypes.byref(rank))
    if rc != 0:
        raise RuntimeError("oracle_pluq stack overflow")
    return dict(
        U=U, L=np.asfortranarray(L[:, :m]), L_extra=L[:, m:],
        row_swaps=[(int(rs[2 * i]), int(rs[2 * i + 1])) for i in range(nr.value)],
        col_swaps=[(int(cs[2 * i]), int(cs[2 * i + 1])) for i in range(nc.value)],
        rank=int(rank.value),
    )


def lu(A, p: int, lib_=None) -> dict:
    """PLE: A[perm-1, :] = L @ U mod p.  Returns dict(U, L, perm, pivcols, rank)."""
    A = _as_f64(A)
    m, n = A.shape
    U = np.zeros((m, n), dtype=np.int64, order="F")
    L = np.zeros((m, m), dtype=np.int64, order="F")
    perm = np.zeros(max(m, 1), dtype=np.int64)
    piv = np.zeros(max(min(m, n), 1), dtype=np.int64)
    rank = _i64()
    (lib_ or lib()).oracle_lu(m, n, _ptr(A), max(m, 1), p, _ptr(U), max(m, 1), _ptr(L), max(m, 1),
                    _ptr(perm), _ptr(piv), ctypes.byref(rank))
    r = int(rank.value)
    return dict(U=U, L=L, perm=perm[:m].copy(), pivcols=piv[:r].copy(), rank=r)


def rref(A, p: int) -> dict:
    """Canonical RREF of A mod p.  Returns dict(R, pivcols (1-based), rank)."""
    A = _as_f64(A)
    m, n = A.shape
    R = np.zeros((m, n), dtype=np.int64, order="F")
    piv = np.zeros(max(min(m, n), 1), dtype=np.int64)
    rank = _i64()
    lib().oracle_rref(m, n, _ptr(A), max(m, 1), p, _ptr(R), max(m, 1), _ptr(piv), ctypes.byref(rank))
    r = int(rank.value)
    return dict(R=R, pivcols=piv[:r].copy(), rank=r)


def inverse(A, p: int):
    """A^{-1} mod p by Gauss-Jordan on [A | I] (the reduced form of an
    invertible A's augmented matrix is [I | A^{-1}]); None when A is singular
    mod p.  For a triangular A this is also its triangular inverse."""
    A = np.asarray(A, dtype=np.int64) % p
    n = A.shape[0]
    if n == 0:
        return np.zeros((0, 0), dtype=np.int64)
    r = rref(np.hstack([A, np.eye(n, dtype=np.int64)]), p)
    if r["rank"] < n or int(r["pivcols"][n - 1]) != n:
        return None
    return r["R"][:, n:].copy()


def is_prime(n: int) -> bool:
    return bool(lib().oracle_is_prime(n))


def largest_primes_below(bound: int, count: int) -> List[int]:
    out = []
    x = bound - 1
    while len(out) < count and x >= 2:
        if is_prime(x):
            out.append(x)
        x -= 1
    return out


# ---------------------------------------------------------------------------
# permutation stacks (src/CuModMatrix/rref_lu_pluq/permutations.jl:11-157)
# ---------------------------------------------------------------------------

def apply_row_perm(stack, X):
    """permutations.jl:73-80/112-125: swap rows (a,b) in stack order."""
    X = np.array(X, copy=True)
    for a, b in stack:
        X[[a - 1, b - 1], :] = X[[b - 1, a - 1], :]
    return X


def apply_row_inv_perm(stack, X):
    """permutations.jl:82-89: the stack applied in reverse order."""
    return apply_row_perm(list(reversed(stack)), X)


def apply_col_perm(stack, X):
    """permutations.jl:11-18/49-62: swap columns (a,b) in stack order."""
    X = np.array(X, copy=True)
    for a, b in stack:
        X[:, [a - 1, b - 1]] = X[:, [b - 1, a - 1]]
    return X


def apply_col_inv_perm(stack, X):
    """permutations.jl:20-27."""
    return apply_col_perm(list(reversed(stack)), X)


def perm_stack_to_matrix(stack, n: int) -> np.ndarray:
    """permutations.jl:141-157 with perm_stack=true: swap rows of I."""
    return apply_row_perm(stack, np.eye(n, dtype=np.int64))


# ---------------------------------------------------------------------------
# independent checks (pure Python / numpy object arithmetic)
# ---------------------------------------------------------------------------

def matmul_bigint(A, B, p: int) -> np.ndarray:
    """Exact product with Python integers, then mod p (brute force)."""
    Ao = np.asarray(A, dtype=object)
    Bo = np.asarray(B, dtype=object)
    C = Ao.dot(Bo) if Ao.size and Bo.size else np.zeros((Ao.shape[0], Bo.shape[1]), dtype=object)
    return np.asarray(C % p, dtype=np.int64)


def _mod_matvec(M: np.ndarray, x: np.ndarray, p: int) -> np.ndarray:
    """M @ x mod p for int64 M with entries in [0,p) and p < 2^31, exact."""
    if M.shape[1] == 0:
        return np.zeros(M.shape[0], dtype=np.int64)
    # split columns so that each partial dot stays below 2^63
    step = max(1, (2 ** 62) // max(1, (p - 1) * (p - 1)))
    acc = np.zeros(M.shape[0], dtype=np.int64)
    for s in range(0, M.shape[1], step):
        acc = (acc + (M[:, s:s + step].astype(np.int64) @ x[s:s + step].astype(np.int64)) % p) % p
    return acc


def freivalds_equal(lhs_apply, rhs_apply, ncols: int, p: int, trials: int = 3, seed: int = 12345) -> bool:
    """Randomised check that two linear maps agree: lhs(x) == rhs(x) mod p for
    random x.  A mismatch is missed with probability <= (1/p)^trials."""
    rng = np.random.default_rng(seed)
    for _ in range(trials):
        x = rng.integers(0, p, size=ncols, dtype=np.int64)
        if not np.array_equal(lhs_apply(x) % p, rhs_apply(x) % p):
            return False
    return True


def check_pluq(A, p: int, U, L, row_swaps, col_swaps, rank: int, exact: bool = True) -> List[str]:
    """Invariants of the reference's PLUQ output (pluq_kernels.jl:46-157):
    P*A*Q == L*U mod p with P = row stack applied to rows, Q = column stack
    applied to columns; U row echelon with unit pivots at (k,k) for k < rank
    and zero rows below rank; L lower triangular with L[k,k] != 0 for k < rank
    and zero columns beyond rank.  Returns the list of violated properties."""
    A = np.asarray(A, dtype=np.int64)
    U = np.asarray(U, dtype=np.int64)
    L = np.asarray(L, dtype=np.int64)
    m, n = A.shape
    errs = []
    if U.shape != (m, n) or L.shape != (m, m):
        return ["shape"]
    r = rank
    for k in range(r):
        if U[k, k] != 1:
            errs.append(f"U[{k},{k}] != 1")
            break
        if L[k, k] == 0:
            errs.append(f"L[{k},{k}] == 0")
            break
    if np.any(np.tril(U, -1)):
        errs.append("U not upper triangular")
    if r < m and np.any(U[r:, :]):
        errs.append("U rows beyond rank nonzero")
    if np.any(np.triu(L, 1)):
        errs.append("L not lower triangular")
    if r < m and np.any(L[:, r:]):
        errs.append("L columns beyond rank nonzero")
    PAQ = apply_col_perm(col_swaps, apply_row_perm(row_swaps, A))
    if exact:
        LU = matmul_bigint(L, U, p) if m * n * m <= 4_000_000 else _blocked_mod_matmul(L, U, p)
        if not np.array_equal(PAQ % p, LU):
            errs.append("P*A*Q != L*U")
    else:
        ok = freivalds_equal(lambda x: _mod_matvec(PAQ, x, p),
                             lambda x: _mod_matvec(L, _mod_matvec(U, x, p), p), n, p)
        if not ok:
            errs.append("P*A*Q != L*U (Freivalds)")
    return errs


def _blocked_mod_matmul(A, B, p: int) -> np.ndarray:
    """Exact A @ B mod p through the C oracle (for medium sizes)."""
    return gemm(np.mod(A, p), np.mod(B, p), p)


def is_rref(R, p: int) -> Tuple[bool, List[int]]:
    """Check the reduced-row-echelon-form conditions; returns (ok, pivcols 1-based)."""
    R = np.asarray(R, dtype=np.int64)
    m, n = R.shape
    piv = []
    last = -1
    zero_seen = False
    for i in range(m):
        nz = np.nonzero(R[i])[0]
        if nz.size == 0:
            zero_seen = True
            continue
        if zero_seen:
            return False, piv
        c = int(nz[0])
        if c <= last or R[i, c] != 1:
            return False, piv
        col = R[:, c]
        if np.count_nonzero(col) != 1:
            return False, piv
        piv.append(c + 1)
        last = c
    return True, piv


def rref_bigint(A, p: int) -> Tuple[np.ndarray, List[int]]:
    """Textbook Gauss-Jordan in Python integers (small inputs only)."""
    M = [[int(x) % p for x in row] for row in np.asarray(A)]
    m = len(M)
    n = len(M[0]) if m else 0
    r = 0
    piv = []
    for c in range(n):
        at = next((i for i in range(r, m) if M[i][c]), None)
        if at is None:
            continue
        M[r], M[at] = M[at], M[r]
        inv = pow(M[r][c], p - 2, p) if _is_prime_small(p) else _egcd_inv(M[r][c], p)
        M[r] = [(x * inv) % p for x in M[r]]
        for i in range(m):
            if i != r and M[i][c]:
                f = M[i][c]
                M[i] = [(a - f * b) % p for a, b in zip(M[i], M[r])]
        piv.append(c + 1)
        r += 1
        if r == m:
            break
    return np.array(M, dtype=np.int64).reshape(m, n), piv


def _egcd_inv(a: int, p: int) -> int:
    return pow(a, -1, p)


def _is_prime_small(p: int) -> bool:
    if p < 2:
        return False
    i = 2
    while i * i <= p:
        if p % i == 0:
            return False
        i += 1
    return True


# ---------------------------------------------------------------------------
# the pivot rule, independently of the C restatement
# ---------------------------------------------------------------------------

def findmax_first(values) -> Tuple[int, int]:
    """Base.findmax semantics (value, 0-based index of the FIRST maximal
    element), which pluq_kernels.jl:188-189 relies on.  Implemented with
    numpy.argmax, whose documented tie rule is the same: "in case of multiple
    occurrences of the maximum values, the indices corresponding to the first
    occurrence are returned"."""
    v = np.asarray(values, dtype=np.int64)
    i = int(np.argmax(v))
    return int(v[i]), i


def pluq_restated(A, p: int, mode: int = PLUQ_OPTC) -> dict:
    """A second, pure-Python restatement of pluq_gpu_kernel
    (pluq_kernels.jl:46-157), written from the reference in Python integers
    with the pivot search delegated to findmax_first.  mode PLUQ_SRC follows
    the reference loop as written (zero column: swap with the last column,
    col += 1, L indexed by col); PLUQ_OPTC applies the corrected zero-column
    handling (swap with a shrinking last column and retry; all-zero block:
    the tail of swaps (c, n)) and indexes L by the pivot count.  Small inputs
    only: it exists to pin the C restatement's pivot choices."""
    M = [[int(x) % p for x in row] for row in np.asarray(A, dtype=np.int64).tolist()]
    m = len(M)
    n = len(M[0]) if m else 0
    lcols = max(m, n)
    Lw = [[0] * lcols for _ in range(m)]
    SR, SC = [], []
    row = col = 0
    last = n - 1
    # swap_rows(d_L, ...) is launched over 32*cld(rows,32) columns (:254,:282)
    lswap = min(lcols, 32 * ((m + 31) // 32))
    while row < m and col < n:
        v, i = findmax_first([M[r][col] for r in range(row, m)])
        if v == 0:
            if mode == PLUQ_SRC:
                # find_pivot :191-198: swap with column `cols`, push, col += 1
                for r in range(m):
                    M[r][col], M[r][n - 1] = M[r][n - 1], M[r][col]
                SC.append((col + 1, n))
                col += 1
                continue
            if all(M[r][c] == 0 for r in range(row, m) for c in range(col, n)):
                for c in range(col, n):
                    for r in range(m):
                        M[r][c], M[r][n - 1] = M[r][n - 1], M[r][c]
                    SC.append((c + 1, n))
                break
            for r in range(m):
                M[r][col], M[r][last] = M[r][last], M[r][col]
            SC.append((col + 1, last + 1))
            last -= 1
            continue
        pr = row + i
        inv = pow(v, -1, p)
        # swap_rows_and_mod(d_A, prow, row, ...) :307-320
        old = M[pr]
        M[pr] = M[row]
        M[row] = [(x * inv) % p for x in old]
        if pr != row:
            for c in range(lswap):
                Lw[row][c], Lw[pr][c] = Lw[pr][c], Lw[row][c]
            SR.append((row + 1, pr + 1))
        lc = col if mode == PLUQ_SRC else row
        # move_and_zero_out :340-351
        Lw[row][lc] = v
        mult = [0] * m
        for r in range(row + 1, m):
            Lw[r][lc] = M[r][col]
            mult[r] = M[r][col]
            M[r][col] = 0
        # update_sub_matrix_kernel :411-440
        for r in range(row + 1, m):
            f = p - mult[r]
            for c in range(col + 1, n):
                M[r][c] = (M[r][c] + f * M[row][c]) % p
        row += 1
        col += 1
    U = np.array(M, dtype=np.int64).reshape(m, n)
    Lfull = np.array(Lw, dtype=np.int64).reshape(m, lcols)
    return dict(U=U, L=Lfull[:, :m].copy(), L_extra=Lfull[:, m:].copy(), row_swaps=SR, col_swaps=SC, rank=row)


def lu_restated(A, p: int) -> dict:
    """Pure-Python PLE with the same pivot rule (findmax_first): a zero column
    is skipped without a column swap; returns dict(U, L, perm, pivcols, rank)
    like the C lu (intent of test/Experiments/rref_gpu_type.jl:60-103)."""
    M = [[int(x) % p for x in row] for row in np.asarray(A, dtype=np.int64).tolist()]
    m = len(M)
    n = len(M[0]) if m else 0
    Lw = [[0] * m for _ in range(m)]
    perm = list(range(1, m + 1))
    piv = []
    row = col = 0
    while row < m and col < n:
        v, i = findmax_first([M[r][col] for r in range(row, m)])
        if v == 0:
            col += 1
            continue
        pr = row + i
        inv = pow(v, -1, p)
        old = M[pr]
        M[pr] = M[row]
        M[row] = [(x * inv) % p for x in old]
        if pr != row:
            Lw[row], Lw[pr] = Lw[pr], Lw[row]
            perm[row], perm[pr] = perm[pr], perm[row]
        Lw[row][row] = v
        for r in range(row + 1, m):
            f = M[r][col]
            Lw[r][row] = f
            M[r][col] = 0
            for c in range(col + 1, n):
                M[r][c] = (M[r][c] - f * M[row][c]) % p
        piv.append(col + 1)
        row += 1
        col += 1
    return dict(U=np.array(M, dtype=np.int64).reshape(m, n), L=np.array(Lw, dtype=np.int64).reshape(m, m),
                perm=np.array(perm, dtype=np.int64), pivcols=np.array(piv, dtype=np.int64), rank=row)


def pivot_targets_from_stack(row_swaps, rank: int) -> List[int]:
    """0-based source row of each pivot from a PLUQ row stack: the swap pushed
    at pivot k is (k+1, at+1); a pivot without an entry did not move."""
    at = list(range(rank))
    for a, b in row_swaps:
        at[a - 1] = b - 1
    return at


def pivot_targets_from_perm(perm, rank: int) -> List[int]:
    """0-based source row of each pivot of a PLE from its final row
    permutation.  Pivot k swaps positions k and at_k >= k, and later pivots
    never touch position k again, so the decomposition is unique: replay it
    forwards."""
    perm = [int(x) for x in perm]
    cur = list(range(1, len(perm) + 1))
    where = {v: i for i, v in enumerate(cur)}
    at = []
    for k in range(rank):
        j = where[perm[k]]
        if j < k:
            raise ValueError("perm is not a forward transposition sequence")
        at.append(j)
        vk, vj = cur[k], cur[j]
        cur[k], cur[j] = vj, vk
        where[vj], where[vk] = k, j
    return at


def check_pivot_rule(L, rank: int, at: List[int]) -> List[str]:
    """The pivot rule of pluq_kernels.jl:188-189 (largest residue, FIRST
    position among equal residues), checked on a factorisation's L.

    Column k of L holds the pivot L[k,k] and, below it, the residues of the
    pivot column at step k -- in the FINAL row order.  Undoing the row swaps
    of pivots k+1.. (from the last pivot down) restores the order of step k.
    In that order, with the step's own swap (k, at_k) undone as well, the
    pivot is the first maximum of rows k..m-1: every row in k..at_k-1 is
    strictly smaller, every row after at_k is at most equal."""
    L = np.array(L, dtype=np.int64, copy=True)
    m = L.shape[0]
    errs = []
    for k in range(rank - 1, -1, -1):
        col = L[:, k].copy()
        a = at[k]
        v = col[k]
        # order of step k before its own swap: rows k and a exchanged back
        col[k], col[a] = col[a], col[k]
        seg = col[k:]
        if v <= 0 or seg[a - k] != v:
            errs.append(f"pivot {k}: value bookkeeping")
            break
        if np.any(seg > v):
            errs.append(f"pivot {k}: a larger residue was below the pivot")
            break
        if np.any(seg[:a - k] == v):
            errs.append(f"pivot {k}: an earlier row had an equal residue (tie not broken to the first)")
            break
        # undo this pivot's swap on the rows of L for the next (earlier) step
        if a != k:
            L[[k, a], :] = L[[a, k], :]
    return errs
