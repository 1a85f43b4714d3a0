"""Synthetic inputs of the batched workloads, shared by the benchmark and the
tests so that both run the same instances.

Host-side numpy only (no device work, no library call): the inputs are drawn
from numpy's PCG64 generator with a fixed seed.

* Batched products: ``count`` independent C_i = A_i * B_i mod p_i with m, n,
  k uniform in [16, 256] and p_i uniform over the 64 largest primes below
  2^23; entries uniform on [0, p_i).  All operands live in one flat,
  column-major float64 arena: A_i (m x k), B_i (k x n), C_i (m x n),
  back to back.
* Batched eliminations: ``count`` matrices with rows and columns uniform in
  [8, 128], p uniform over {2, 3, 7, 251, 65521}, entries uniform on
  [0, p).  Every fourth instance is made rank-deficient: a random subset of
  its rows is zeroed, then it is left-multiplied by a random matrix that is
  invertible mod p (redrawn until it is).
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

ELIM_PRIMES = (2, 3, 7, 251, 65521)


def _is_prime(x: int) -> bool:
    """Deterministic Miller-Rabin for x < 3.3e24 (the first 12 prime bases)."""
    if x < 2:
        return False
    bases = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)
    for b in bases:
        if x % b == 0:
            return x == b
    d, s = x - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in bases:
        y = pow(a, d, x)
        if y in (1, x - 1):
            continue
        for _ in range(s - 1):
            y = y * y % x
            if y == x - 1:
                break
        else:
            return False
    return True


def largest_primes_below(bound: int, count: int) -> List[int]:
    out, x = [], bound - 1
    while len(out) < count and x >= 2:
        if _is_prime(x):
            out.append(x)
        x -= 1
    return out


class GemmBatch:
    """The batched-product workload: shapes, primes, arena offsets (in
    elements) and the host arena with A_i and B_i filled and C_i zero."""

    def __init__(self, count: int = 10000, seed: int = 3, lo: int = 16, hi: int = 256):
        rng = np.random.default_rng(seed)
        primes = largest_primes_below(2 ** 23, 64)
        self.shapes = rng.integers(lo, hi + 1, size=(count, 3)).astype(np.int64)  # (m, n, k)
        self.primes = np.array([primes[i] for i in rng.integers(0, 64, size=count)], dtype=np.int64)
        m, n, k = self.shapes[:, 0], self.shapes[:, 1], self.shapes[:, 2]
        sizes = m * k + k * n + m * n
        self.offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        self.elems = int(sizes.sum())
        self.arena = np.zeros(self.elems, dtype=np.float64)
        for i in range(count):
            mi, ni, ki = (int(x) for x in self.shapes[i])
            o = int(self.offsets[i])
            p = int(self.primes[i])
            self.arena[o:o + mi * ki + ki * ni] = rng.integers(0, p, size=mi * ki + ki * ni)
        self.ops = int((2 * m * n * k).sum())
        self.count = count

    def a_off(self, i: int) -> int:
        return int(self.offsets[i])

    def b_off(self, i: int) -> int:
        m, _, k = self.shapes[i]
        return int(self.offsets[i] + m * k)

    def c_off(self, i: int) -> int:
        m, n, k = self.shapes[i]
        return int(self.offsets[i] + m * k + k * n)

    def operands(self, i: int) -> Tuple[np.ndarray, np.ndarray, int]:
        """(A_i, B_i, p_i) as int64 matrices."""
        m, n, k = (int(x) for x in self.shapes[i])
        a, b = self.a_off(i), self.b_off(i)
        A = self.arena[a:a + m * k].reshape(k, m).T.astype(np.int64)
        B = self.arena[b:b + k * n].reshape(n, k).T.astype(np.int64)
        return A, B, int(self.primes[i])


def rank_mod_p(M: np.ndarray, p: int) -> int:
    """Rank of a small integer matrix mod p by row reduction (numpy int64;
    p <= 65521 keeps every product below 2^63)."""
    W = np.array(M, dtype=np.int64) % p
    rows, cols = W.shape
    r = 0
    for c in range(cols):
        if r == rows:
            break
        nz = np.nonzero(W[r:, c])[0]
        if nz.size == 0:
            continue
        piv = r + int(nz[0])
        if piv != r:
            W[[r, piv]] = W[[piv, r]]
        inv = pow(int(W[r, c]), -1, p)
        W[r] = (W[r] * inv) % p
        below = W[r + 1:, c].copy()
        W[r + 1:] = (W[r + 1:] - below[:, None] * W[r][None, :]) % p
        r += 1
    return r


def _matmul_mod(A: np.ndarray, B: np.ndarray, p: int) -> np.ndarray:
    """A @ B mod p for entries in [0, p), p <= 65521 and inner size <= 128
    (every partial sum stays below 128 * 65520^2 < 2^63)."""
    return (A.astype(np.int64) @ B.astype(np.int64)) % p


def elim_instances(count: int = 4096, seed: int = 4, lo: int = 8, hi: int = 128) -> List[Tuple[np.ndarray, int]]:
    """The batched-elimination workload: a list of (A int64 m x n, p)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        m, n = (int(x) for x in rng.integers(lo, hi + 1, size=2))
        p = ELIM_PRIMES[int(rng.integers(0, len(ELIM_PRIMES)))]
        A = rng.integers(0, p, size=(m, n)).astype(np.int64)
        if i % 4 == 0:
            zero = rng.choice(m, size=int(rng.integers(1, m)), replace=False)
            A[zero, :] = 0
            while True:
                S = rng.integers(0, p, size=(m, m)).astype(np.int64)
                if rank_mod_p(S, p) == m:
                    break
            A = _matmul_mod(S, A, p)
        out.append((A, p))
    return out


# ---------------------------------------------------------------------------
# device layouts of the two workloads (job tables over one arena each)
# ---------------------------------------------------------------------------

def gemm_job_table(batch: GemmBatch, base_ptr: int):
    """ctypes array of modp_gemm_job over a device copy of ``batch.arena``
    starting at ``base_ptr``."""
    from ._lib import GemmJob
    jobs = (GemmJob * batch.count)()
    for i in range(batch.count):
        m, n, k = (int(x) for x in batch.shapes[i])
        jobs[i] = GemmJob(m, n, k, int(batch.primes[i]), base_ptr + 8 * batch.a_off(i), m,
                          base_ptr + 8 * batch.b_off(i), k, base_ptr + 8 * batch.c_off(i), m)
    return jobs


class ElimArena:
    """Device arena for a batch of eliminations: inputs A_i, outputs U_i (R_i
    for rref) and L_i (float64, column-major, ld = rows), and per instance
    int64 row information (max(2 min(m,n), m) entries), column information
    (max(4 n, min(m,n))) and three counts, with the job table over it."""

    def __init__(self, ctx, instances):
        from ._lib import ElimJob
        self.ctx = ctx
        self.instances = instances
        self.sizes = [(int(A.shape[0]), int(A.shape[1])) for A, _ in instances]
        self.nA = sum(m * n for m, n in self.sizes)
        self.nL = sum(m * m for m, _ in self.sizes)
        self.ri_len = [max(2 * min(m, n), m) for m, n in self.sizes]
        self.ci_len = [max(4 * n, min(m, n)) for m, n in self.sizes]
        self.nI = sum(r + c + 3 for r, c in zip(self.ri_len, self.ci_len))
        self.fbuf = ctx.malloc(8 * (2 * self.nA + self.nL))
        self.ibuf = ctx.malloc(8 * self.nI)
        host = np.zeros(self.nA, dtype=np.float64)
        self.jobs = (ElimJob * len(instances))()
        self.offs = []
        oa = ol = oi = 0
        for i, ((A, p), (m, n)) in enumerate(zip(instances, self.sizes)):
            host[oa:oa + m * n] = A.astype(np.float64).ravel(order="F")
            a = self.fbuf.ptr + 8 * oa
            u = self.fbuf.ptr + 8 * (self.nA + oa)
            l_ = self.fbuf.ptr + 8 * (2 * self.nA + ol)
            ri = self.ibuf.ptr + 8 * oi
            ci = ri + 8 * self.ri_len[i]
            cn = ci + 8 * self.ci_len[i]
            self.jobs[i] = ElimJob(m, n, int(p), a, m, u, m, l_, m, ri, ci, cn)
            self.offs.append((oa, ol, oi))
            oa += m * n
            ol += m * m
            oi += self.ri_len[i] + self.ci_len[i] + 3
        self.fbuf.upload(host)
        self.bytes_moved = 8 * (2 * self.nA + self.nL)

    def clear_outputs(self):
        """Zero U, L and the index outputs (the inputs stay)."""
        lib, h = self.ctx.lib, self.ctx.h
        self.ctx.check(lib.modp_memset_zero(h, self.fbuf.ptr + 8 * self.nA, 8 * (self.nA + self.nL)), "memset")
        self.ibuf.zero()

    def download(self):
        """All outputs: a list of dicts (U, L, row, col, counts) per instance."""
        f = np.empty(2 * self.nA + self.nL, dtype=np.float64)
        self.fbuf.download(f)
        info = np.empty(self.nI, dtype=np.int64)
        self.ibuf.download(info)
        out = []
        for i, (m, n) in enumerate(self.sizes):
            oa, ol, oi = self.offs[i]
            U = f[self.nA + oa:self.nA + oa + m * n].reshape(n, m).T.astype(np.int64)
            L = f[2 * self.nA + ol:2 * self.nA + ol + m * m].reshape(m, m).T.astype(np.int64)
            r = info[oi:oi + self.ri_len[i]]
            c = info[oi + self.ri_len[i]:oi + self.ri_len[i] + self.ci_len[i]]
            cnt = info[oi + self.ri_len[i] + self.ci_len[i]:oi + self.ri_len[i] + self.ci_len[i] + 3]
            out.append(dict(U=U, L=L, row=r, col=c, counts=cnt))
        return out
