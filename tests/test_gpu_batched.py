"""Batched multi-modular forms and device permutations on the GPU."""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle as O
from gpufinitefieldmatrices import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))


def test_permutations_reference_fixture(gffm):
    g = gffm
    f = GOLD["permutation"]
    st = [tuple(x) for x in f["stack"]]
    A = g.CuModMatrix(np.array(f["A"]), f["p"])
    g.apply_col_perm_(st, A)
    assert g.Array(A).astype(int).tolist() == f["col_perm"]
    g.apply_col_inv_perm_(st, A)
    assert g.Array(A).astype(int).tolist() == f["A"]
    A = g.CuModMatrix(np.array(f["A"]), f["p"])
    g.apply_row_perm_(st, A)
    assert g.Array(A).astype(int).tolist() == f["row_perm"]
    g.apply_row_inv_perm_(st, A)
    assert g.Array(A).astype(int).tolist() == f["A"]


@pytest.mark.parametrize("et", [np.float64, np.float32])
@pytest.mark.parametrize("nswaps", [3, 9, 200, 1000])
def test_long_swap_stacks_match_oracle(gffm, nswaps, et):
    # stacks longer than a few swaps are composed into one permutation on the
    # host: repeated indices, self-swaps and cycles must give the sequential
    # result, for rows and columns, forwards and inverse
    g = gffm
    p = 251
    rng = np.random.default_rng(nswaps)
    Ah = rng.integers(0, p, size=(90, 70))
    for rowwise, lim in ((True, 90), (False, 70)):
        st = [(int(a), int(b)) for a, b in rng.integers(1, lim + 1, size=(nswaps, 2))]
        st[0] = (st[0][0], st[0][0])
        fwd = O.apply_row_perm if rowwise else O.apply_col_perm
        inv = O.apply_row_inv_perm if rowwise else O.apply_col_inv_perm
        for f_dev, f_ref in ((g.apply_row_perm_ if rowwise else g.apply_col_perm_, fwd),
                             (g.apply_row_inv_perm_ if rowwise else g.apply_col_inv_perm_, inv)):
            A = g.CuModMatrix(Ah, p, elem_type=et)
            f_dev(st, A)
            assert np.array_equal(g.Array(A).astype(np.int64), f_ref(st, Ah))


def test_pluq_identity_through_permutations(gffm):
    # P*A*Q == L*U using the device permutation kernels and the device GEMM
    g = gffm
    p = 65521
    Ah = O.draw_matrix(70, 50, p, 3)
    Ah[:, 10] = 0
    A = g.CuModMatrix(Ah, p, elem_type=np.float64)
    U, L, P, Q = g.pluq_gpu_kernel(A)
    B = g.copy(A)
    g.apply_row_perm_(P, B)
    g.apply_col_perm_(Q, B)
    assert np.array_equal(g.Array(B), g.Array(L * U))


def gemm_jobs(g, count, seed, lo=16, hi=256):
    rng = np.random.default_rng(seed)
    primes = O.largest_primes_below(2 ** 23, 64)
    jobs, host = [], []
    for i in range(count):
        m, n, k = (int(x) for x in rng.integers(lo, hi + 1, size=3))
        p = primes[int(rng.integers(0, 64))]
        Ah = O.draw_matrix(m, k, p, seed=seed * 100000 + i)
        Bh = O.draw_matrix(k, n, p, seed=seed * 100000 + i, offset=m * k)
        A = g.CuModMatrix(Ah, p, elem_type=np.float64)
        B = g.CuModMatrix(Bh, p, elem_type=np.float64)
        C = g.zeros(np.float64, m, n, p)
        jobs.append((A, B, C, p))
        host.append((Ah, Bh, p))
    return jobs, host


def test_batched_gemm_vs_oracle(gffm):
    g = gffm
    jobs, host = gemm_jobs(g, 150, seed=3)
    plan = g.GemmBatchPlan(jobs)
    for rep in range(3):
        # replays of the captured graph, with eager work in between
        for _, _, C, _ in jobs:
            C.data.zero()
        plan.run()
        g.context().synchronize()
        for (A, B, C, p), (Ah, Bh, _) in zip(jobs, host):
            assert np.array_equal(g.Array(C).astype(np.int64), O.gemm(Ah, Bh, p))


def test_batched_gemm_edge_shapes(gffm):
    g = gffm
    jobs, host = gemm_jobs(g, 40, seed=9, lo=1, hi=70)
    g.gemm_batched(jobs)
    for (A, B, C, p), (Ah, Bh, _) in zip(jobs, host):
        assert np.array_equal(g.Array(C).astype(np.int64), O.gemm(Ah, Bh, p))


class ElimBatch:
    def __init__(self, g, inst, op):
        self.g, self.inst, self.op = g, inst, op
        ctx = g.context()
        self.keep = []
        jobs = []
        for A, p in inst:
            m, n = A.shape
            dA = g.CuModMatrix(A, p, elem_type=np.float64)
            dU = g.zeros(np.float64, m, n, p)
            dL = g.zeros(np.float64, m, m, p)
            ri = ctx.malloc(8 * max(2 * min(m, n), m, 1))
            ci = ctx.malloc(8 * max(4 * n, 1))
            cnt = ctx.malloc(8 * 3)
            self.keep.append((dA, dU, dL, ri, ci, cnt))
            jobs.append(g.ElimJob(m, n, p, dA.ptr, dA.ld, dU.ptr, dU.ld, dL.ptr, dL.ld, ri.ptr, ci.ptr, cnt.ptr))
        self.jobs = jobs

    def run(self):
        self.g.elim_batched(self.op, self.jobs)

    def result(self, i):
        g = self.g
        dA, dU, dL, ri, ci, cnt = self.keep[i]
        m, n = self.inst[i][0].shape
        c = np.zeros(3, dtype=np.int64)
        cnt.download(c)
        rb = np.zeros(ri.nbytes // 8, dtype=np.int64)
        ri.download(rb)
        cb = np.zeros(ci.nbytes // 8, dtype=np.int64)
        ci.download(cb)
        return dict(U=g.Array(dU).astype(np.int64), L=g.Array(dL).astype(np.int64), counts=c, row=rb, col=cb)


@pytest.mark.parametrize("op", ["pluq", "lu", "rref"])
def test_batched_elim_vs_oracle(gffm, op):
    g = gffm
    inst = W.elim_instances(300, seed=4)
    code = {"pluq": 0, "lu": 1, "rref": 2}[op]
    b = ElimBatch(g, inst, code)
    b.run()
    for i, (A, p) in enumerate(inst):
        r = b.result(i)
        if op == "pluq":
            ref = O.pluq(A, p, O.PLUQ_OPTC)
            nr, nc, rank = r["counts"]
            assert rank == ref["rank"], i
            assert [tuple(x) for x in r["row"][:2 * nr].reshape(-1, 2).tolist()] == ref["row_swaps"]
            assert [tuple(x) for x in r["col"][:2 * nc].reshape(-1, 2).tolist()] == ref["col_swaps"]
            assert np.array_equal(r["U"], ref["U"]) and np.array_equal(r["L"], ref["L"]), i
        elif op == "lu":
            ref = O.lu(A, p)
            rank = r["counts"][2]
            assert rank == ref["rank"]
            assert r["row"][:A.shape[0]].tolist() == ref["perm"].tolist()
            assert r["col"][:rank].tolist() == ref["pivcols"].tolist()
            assert np.array_equal(r["U"], ref["U"]) and np.array_equal(r["L"], ref["L"])
        else:
            ref = O.rref(A, p)
            rank = r["counts"][2]
            assert rank == ref["rank"]
            assert r["row"][:rank].tolist() == ref["pivcols"].tolist()
            assert np.array_equal(r["U"], ref["R"])


# ---------------------------------------------------------------- full batches
# The whole batched workloads that the benchmark times, every instance
# compared with the oracle.

def test_full_gemm_batch_vs_oracle(gffm):
    # 10,000 products, m, n, k ~ U[16, 256], each under one of the 64 largest
    # primes below 2^23 (seed 3); the graph is replayed and every replay's
    # outputs are compared
    g = gffm
    ctx = g.context()
    wb = W.GemmBatch(10000, seed=3)
    print(f"full gemm batch: {wb.count} products, {wb.elems * 8 / 1e9:.2f} GB arena", flush=True)
    buf = ctx.malloc(8 * wb.elems)
    buf.upload(wb.arena)
    jobs = W.gemm_job_table(wb, buf.ptr)
    import ctypes
    h = ctypes.c_void_p()
    ctx.check(ctx.lib.modp_gemm_batched_plan(ctx.h, 0, jobs, wb.count, ctypes.byref(h)))
    try:
        refs = []
        for i in range(wb.count):
            A, B, p = wb.operands(i)
            refs.append(O.gemm(A, B, p))
        out = np.empty(wb.elems, dtype=np.float64)
        for rep in range(2):
            ctx.check(ctx.lib.modp_gemm_batched_run(ctx.h, h))
            ctx.synchronize()
            buf.download(out)
            bad = []
            for i in range(wb.count):
                m, n, _ = (int(x) for x in wb.shapes[i])
                c = wb.c_off(i)
                C = out[c:c + m * n].reshape(n, m).T.astype(np.int64)
                if not np.array_equal(C, refs[i]):
                    bad.append(i)
            assert bad == [], f"replay {rep}: {len(bad)} products differ, first {bad[:5]}"
            # eager work between replays: clear every C
            buf.upload(wb.arena)
    finally:
        ctx.check(ctx.lib.modp_gemm_batched_destroy(ctx.h, h))


@pytest.fixture(scope="module")
def full_elim_batch():
    return W.elim_instances(4096, seed=4)


def _elim_mismatches(op, res, inst):
    bad = []
    for i, ((A, p), r) in enumerate(zip(inst, res)):
        m, n = A.shape
        if op == "pluq":
            ref = O.pluq(A, p, O.PLUQ_OPTC)
            nr, nc, rank = (int(x) for x in r["counts"])
            ok = (rank == ref["rank"]
                  and [tuple(x) for x in r["row"][:2 * nr].reshape(-1, 2).tolist()] == ref["row_swaps"]
                  and [tuple(x) for x in r["col"][:2 * nc].reshape(-1, 2).tolist()] == ref["col_swaps"]
                  and np.array_equal(r["U"], ref["U"]) and np.array_equal(r["L"], ref["L"]))
        elif op == "lu":
            ref = O.lu(A, p)
            rank = int(r["counts"][2])
            ok = (rank == ref["rank"] and r["row"][:m].tolist() == ref["perm"].tolist()
                  and r["col"][:rank].tolist() == ref["pivcols"].tolist()
                  and np.array_equal(r["U"], ref["U"]) and np.array_equal(r["L"], ref["L"]))
        else:
            ref = O.rref(A, p)
            rank = int(r["counts"][2])
            ok = (rank == ref["rank"] and r["row"][:rank].tolist() == ref["pivcols"].tolist()
                  and np.array_equal(r["U"], ref["R"]))
        if not ok:
            bad.append(i)
    return bad


@pytest.mark.parametrize("op", ["pluq", "lu", "rref"])
def test_full_elim_batch_vs_oracle(gffm, full_elim_batch, op):
    # 4096 eliminations, rows and columns ~ U[8, 128], p in {2, 3, 7, 251,
    # 65521}, every fourth rank-deficient (seed 4): every instance's outputs
    # equal the oracle's
    g = gffm
    ctx = g.context()
    print(f"full elimination batch: {op}", flush=True)
    arena = W.ElimArena(ctx, full_elim_batch)
    code = {"pluq": 0, "lu": 1, "rref": 2}[op]
    for rep in range(2):
        arena.clear_outputs()
        g.elim_batched(code, list(arena.jobs), ctx=ctx)
        ctx.synchronize()
        bad = _elim_mismatches(op, arena.download(), full_elim_batch)
        assert bad == [], f"run {rep}: {len(bad)} instances differ, first {bad[:5]}"


def test_elim_plan_replays_vs_oracle(gffm):
    # a captured elimination plan replayed three times with eager work
    # between the replays (outputs cleared, a product run); every replay
    # equals the oracle
    g = gffm
    ctx = g.context()
    inst = W.elim_instances(400, seed=11)
    for op, code in (("pluq", 0), ("lu", 1), ("rref", 2)):
        arena = W.ElimArena(ctx, inst)
        plan = g.ElimBatchPlan(code, list(arena.jobs), ctx=ctx)
        for rep in range(3):
            arena.clear_outputs()
            X = g.rand(np.float64, 64, 64, 251, seed=rep)
            _ = X * X
            plan.run()
            ctx.synchronize()
            bad = _elim_mismatches(op, arena.download(), inst)
            assert bad == [], f"{op} replay {rep}: {len(bad)} instances differ, first {bad[:5]}"


@pytest.mark.parametrize("op", ["pluq", "lu", "rref"])
def test_elim_batch_large_primes_and_edges(gffm, op):
    # moduli above 2^16 (32-bit residues in LDS, 128 x 128 slots above
    # 64 KiB) mixed with small ones, and edge shapes
    g = gffm
    ctx = g.context()
    rng = np.random.default_rng(12)
    primes = [2, 65521, 65537, 8388593, 2 ** 26 - 5]
    shapes = [(1, 1), (1, 9), (9, 1), (128, 128), (128, 3), (3, 128), (127, 128), (64, 65), (65, 64)]
    inst = []
    for i in range(60):
        m, n = shapes[i % len(shapes)] if i < 2 * len(shapes) else (int(x) for x in rng.integers(1, 129, size=2))
        p = primes[i % len(primes)]
        A = rng.integers(0, p, size=(m, n))
        if i % 3 == 0 and m > 1:
            A[rng.choice(m, size=m // 2, replace=False), :] = 0
        if i % 5 == 0 and n > 1:
            A[:, rng.integers(0, n)] = 0
        inst.append((A.astype(np.int64), p))
    arena = W.ElimArena(ctx, inst)
    code = {"pluq": 0, "lu": 1, "rref": 2}[op]
    g.elim_batched(code, list(arena.jobs), ctx=ctx)
    bad = _elim_mismatches(op, arena.download(), inst)
    assert bad == [], f"{len(bad)} instances differ, first {bad[:5]}"


def test_elim_batch_not_prime_reports_minus_one(gffm):
    g = gffm
    ctx = g.context()
    A = np.array([[2, 1], [1, 1]], dtype=np.int64)  # pivot 2 has no inverse mod 4
    arena = W.ElimArena(ctx, [(A, 4), (np.array([[1, 2], [3, 4]], dtype=np.int64), 5)])
    g.elim_batched(0, list(arena.jobs), ctx=ctx)
    res = arena.download()
    assert int(res[0]["counts"][2]) == -1 and int(res[1]["counts"][2]) == 2
