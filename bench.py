"""Benchmark: mod-p GEMM at n = 8192, p = 2^26 - 5 in the int8 multi-modular
MFMA form (the form the library picks there) against the int8 digit form and
the plain FP64-MFMA form (all bit-exact), plus PLUQ seconds at 4096^2 mod
65521 and the other workloads of the hot path (1024^2 plain GEMM, p = 127,
GEMV, batched products, batched eliminations).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--only gemm|pluq|batched]

One "step" is one C = A*B mod p over 8192^3 residues resident in HBM (the
residue splits of A and B, the 9 int8 MFMA products and the Chinese-remainder
recombination all inside the step).  Prints one JSON line.  The 8192^2 operands are generated on the
device by the library's seeded generator; the batched workloads come from
the shared host generator (gpufinitefieldmatrices.workloads) and are
uploaded once before any timing.

The CPU-baseline leg runs the oracle (OpenMP over the host cores) on the
same inputs: it times the full 8192^3 product and keeps it, so the line
reports max |delta| of the timed product against the oracle; it also runs
the oracle on every instance of the two batched workloads and reports how
many device results differ.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "gpufinitefieldmatrices.jl_amd"))

INT8_PEAK_TOPS = 5000.0   # MI355X dense int8 MFMA, TOP/s
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 MFMA
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--only", choices=["all", "gemm", "pluq", "batched"], default="all",
                    help="one section only, for profiling runs: the headline GEMM, the 4096^2 "
                         "eliminations, or the batched forms")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one device copy of 8*n*n bytes (calibrates HBM byte counters)")
    return ap.parse_args()


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def omp_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def time_gemm(g, ctx, A, B, C, form, steps, warmup, kernel_prefix):
    for _ in range(warmup):
        g.mul_(C, A, B, form=form)
    ctx.synchronize()
    ctx.kernel_timing_reset()
    ctx.kernel_timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.event_record(0)
    for _ in range(steps):
        g.mul_(C, A, B, form=form)
    ctx.event_record(1)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ctx.kernel_timing(False)
    ev_ms = ctx.event_elapsed(0, 1)
    ktot, kcnt = ctx.kernel_time(kernel_prefix)
    ctx.kernel_timing_reset()
    return ev_ms / steps, wall / steps * 1e3, (ktot / kcnt if kcnt else None), kcnt


# ---------------------------------------------------------------------------
# batched workloads
# ---------------------------------------------------------------------------

class BatchedGemm:
    """10,000 independent products, m,n,k ~ U[16,256] (seed 3), each under its
    own prime from the 64 largest primes below 2^23, uploaded once."""

    def __init__(self, g, ctx):
        from gpufinitefieldmatrices import workloads as W
        self.g, self.ctx = g, ctx
        self.wb = W.GemmBatch(10000, seed=3)
        self.buf = ctx.malloc(8 * self.wb.elems)
        self.buf.upload(self.wb.arena)
        self.jobs = W.gemm_job_table(self.wb, self.buf.ptr)

    def run(self, reps):
        g, ctx = self.g, self.ctx
        h = ctypes.c_void_p()
        # the plan's tile policy is autotuned (candidate plans built, run and
        # timed before the timed region); the tuner's cache is left as it was
        ctx.autotune_mode(g.TUNE_ON_MISS)
        try:
            ctx.check(ctx.lib.modp_gemm_batched_plan(ctx.h, 0, self.jobs, self.wb.count, ctypes.byref(h)))
        finally:
            ctx.autotune_mode(g.TUNE_OFF)
        ctx.check(ctx.lib.modp_gemm_batched_run(ctx.h, h))
        ctx.synchronize()
        ctx.event_record(2)
        for _ in range(reps):
            ctx.check(ctx.lib.modp_gemm_batched_run(ctx.h, h))
        ctx.event_record(3)
        ms = ctx.event_elapsed(2, 3) / reps
        ctx.check(ctx.lib.modp_gemm_batched_destroy(ctx.h, h))
        wb = self.wb
        return dict(batched_gemm_gops=round(wb.ops / (ms * 1e-3) / 1e9, 2), batched_gemm_ms=round(ms, 4),
                    batched_gemm_hbm_gbs=round(8.0 * wb.elems / (ms * 1e-3) / 1e9, 1))

    def outputs(self):
        out = np.empty(self.wb.elems, dtype=np.float64)
        self.buf.download(out)
        return out


class BatchedElim:
    """4096 independent eliminations, rows/cols ~ U[8,128] (seed 4), prime per
    instance from {2,3,7,251,65521}; every fourth instance rank-deficient."""

    def __init__(self, g, ctx):
        from gpufinitefieldmatrices import workloads as W
        self.g, self.ctx = g, ctx
        self.inst = W.elim_instances(4096, seed=4)
        self.arena = W.ElimArena(ctx, self.inst)

    def run(self, reps=10):
        g, ctx = self.g, self.ctx
        # the plan (sorted job table on the device, captured graph) is built
        # before the timed region; each timed run is one graph launch
        plan = g.ElimBatchPlan(0, list(self.arena.jobs), ctx=ctx)
        plan.run()
        ctx.synchronize()
        ctx.event_record(6)
        for _ in range(reps):
            plan.run()
        ctx.event_record(7)
        ms = ctx.event_elapsed(6, 7) / reps
        sizes = self.arena.sizes
        ops = sum(2 * min(m, n) * m * n for m, n in sizes)
        return dict(batched_pluq_ms=round(ms, 4), batched_pluq_instances=len(sizes),
                    batched_pluq_hbm_gbs=round(self.arena.bytes_moved / (ms * 1e-3) / 1e9, 1),
                    batched_pluq_gops_upper=round(ops / (ms * 1e-3) / 1e9, 2))


# ---------------------------------------------------------------------------
# eliminations at 4096^2
# ---------------------------------------------------------------------------

def _best_of(fn, ctx, reps=3):
    fn()
    ctx.synchronize()
    ts = []
    out = None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        ctx.synchronize()
        ts.append(time.perf_counter() - t0)
    return min(ts), out


def bench_pluq(g, ctx):
    """PLUQ of 4096^2 mod 65521 (seed 2) with both forms of the trailing
    update, and the rank-3000 product (4096x3000)(3000x4096) by PLUQ and RREF
    (best of three runs each)."""
    q, nq = 65521, 4096
    extra = {}
    Aq = g.rand(np.float64, nq, nq, q, seed=2)
    for name, form in (("limb", g.FORM_LIMB_I8), ("plain", g.FORM_PLAIN_F64)):
        t, res = _best_of(lambda: g.pluq(Aq, form=form), ctx)
        extra[f"pluq_sec_{name}"] = round(t, 5)
        extra["pluq_rank"] = res[4]
    extra["pluq_limb_vs_plain"] = round(extra["pluq_sec_plain"] / extra["pluq_sec_limb"], 3)
    X = g.rand(np.float64, nq, 3000, q, seed=2)
    Y = g.rand(np.float64, 3000, nq, q, seed=2, offset=nq * 3000)
    A3 = X * Y
    del X, Y
    t, res = _best_of(lambda: g.pluq(A3, form=g.FORM_LIMB_I8), ctx)
    extra.update(pluq_rank3000_sec=round(t, 5), pluq_rank3000_rank=res[4])
    t, res = _best_of(lambda: g.rref_full(A3, form=g.FORM_LIMB_I8), ctx)
    extra.update(rref_rank3000_sec=round(t, 5))
    del A3
    # A^{-1} of the full-rank 4096^2 matrix (PLUQ, two triangular inverses,
    # permutations, one product)
    t, _ = _best_of(lambda: g.inverse(Aq), ctx)
    extra.update(inverse_4096_sec=round(t, 5))
    del Aq
    return extra


# ---------------------------------------------------------------------------
# CPU baseline (the oracle) and parity counts
# ---------------------------------------------------------------------------

def cpu_baseline(g, n, p, seed, C_dev, batched_gemm, batched_elim):
    """The oracle (C, OpenMP over the host cores) on the same workloads: the
    full n^3 product of the same seeded residues (timed: the baseline), then
    every instance of the batched workloads.  Returns the baseline record and
    the parity counts of the device results against the oracle's."""
    import oracle as O
    A = O.draw_matrix(n, n, p, seed=seed)
    B = O.draw_matrix(n, n, p, seed=seed, offset=n * n)
    t0 = time.perf_counter()
    ref = O.gemm(A, B, p)
    dt = time.perf_counter() - t0
    del A, B
    got = g.Array(C_dev).astype(np.int64)
    parity = {"max_abs_delta_vs_oracle": int(np.abs(got - ref).max())}
    del got, ref
    rec = dict(value=round(2.0 * n ** 3 / dt / 1e9, 3), unit="GOPS", cores=omp_threads(), kind="port",
               cpu_model=cpu_model(),
               sample=f"oracle_gemm full {n}x{n}x{n}, p={p}, seed {seed}: {dt:.2f} s")
    if batched_gemm is not None:
        wb = batched_gemm.wb
        out = batched_gemm.outputs()
        bad = 0
        t0 = time.perf_counter()
        for i in range(wb.count):
            Ah, Bh, pi = wb.operands(i)
            m, nn, _ = (int(x) for x in wb.shapes[i])
            c = wb.c_off(i)
            if not np.array_equal(out[c:c + m * nn].reshape(nn, m).T.astype(np.int64), O.gemm(Ah, Bh, pi)):
                bad += 1
        rec["batched_gemm_oracle_s"] = round(time.perf_counter() - t0, 2)
        parity["batched_gemm_mismatches"] = bad
        parity["batched_gemm_checked"] = wb.count
    if batched_elim is not None:
        res = batched_elim.arena.download()
        bad = 0
        t0 = time.perf_counter()
        for (Ah, pi), r in zip(batched_elim.inst, res):
            ref = O.pluq(Ah, pi, O.PLUQ_OPTC)
            nr, nc, rank = (int(x) for x in r["counts"])
            ok = (rank == ref["rank"]
                  and [tuple(x) for x in r["row"][:2 * nr].reshape(-1, 2).tolist()] == ref["row_swaps"]
                  and [tuple(x) for x in r["col"][:2 * nc].reshape(-1, 2).tolist()] == ref["col_swaps"]
                  and np.array_equal(r["U"], ref["U"]) and np.array_equal(r["L"], ref["L"]))
            bad += 0 if ok else 1
        rec["batched_pluq_oracle_s"] = round(time.perf_counter() - t0, 2)
        parity["batched_pluq_mismatches"] = bad
        parity["batched_pluq_checked"] = len(res)
    return rec, parity


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "the library runs on one GPU; --gpus > 1 is not supported",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import gpufinitefieldmatrices as g
    ctx = g.context()
    if a.only in ("pluq", "batched"):
        # one section, for profiling; the headline GEMM is not run
        if a.only == "pluq":
            out = bench_pluq(g, ctx)
        else:
            out = {**BatchedGemm(g, ctx).run(reps=5), **BatchedElim(g, ctx).run()}
        print(json.dumps({"section": a.only, **out}))
        return
    n = a.n
    p = 2 ** 26 - 5
    ops = 2.0 * n * n * n

    A = g.rand(np.float64, n, n, p, seed=1)
    B = g.rand(np.float64, n, n, p, seed=1, offset=n * n)
    # headline: the multi-modular int8 form (the form FORM_AUTO picks here)
    C_rns = g.zeros(np.float64, n, n, p)
    rns_ms, rns_wall, rns_kms, _ = time_gemm(g, ctx, A, B, C_rns, g.FORM_RNS_I8, a.steps, a.warmup, "gemm_rns")
    extra = {}
    bg = be = None
    if a.only == "all":
        # the base-256 digit form and the plain FP64 form on the same operands
        C_limb = g.zeros(np.float64, n, n, p)
        limb_ms, _, limb_kms, _ = time_gemm(g, ctx, A, B, C_limb, g.FORM_LIMB_I8, a.steps, a.warmup, "gemm_limb")
        C_plain = g.zeros(np.float64, n, n, p)
        plain_ms, _, plain_kms, _ = time_gemm(g, ctx, A, B, C_plain, g.FORM_PLAIN_F64, max(2, a.steps // 2), 1,
                                              "gemm_f64")
        c_rns = g.Array(C_rns)
        maxdelta = float(np.abs(c_rns - g.Array(C_plain)).max())
        maxdelta_limb = float(np.abs(c_rns - g.Array(C_limb)).max())
        del C_plain, C_limb, c_rns
        extra.update(plain_f64_gops=round(ops / (plain_ms * 1e-3) / 1e9, 2),
                     rns_vs_plain=round(plain_ms / rns_ms, 3),
                     limb_digit_gops=round(ops / (limb_ms * 1e-3) / 1e9, 2),
                     limb_digit_ms=round(limb_ms, 4),
                     limb_digit_core_ms=round(limb_kms, 4) if limb_kms else None,
                     limb_digit_vs_plain=round(plain_ms / limb_ms, 3),
                     max_abs_delta_rns_vs_plain=maxdelta, max_abs_delta_rns_vs_limb_digit=maxdelta_limb,
                     plain_core_ms=round(plain_kms, 4) if plain_kms else None,
                     plain_fp64_mfma_frac=round(ops / (plain_kms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4)
                     if plain_kms else None)
        del A, B
        # single-limb prime p = 127
        A7 = g.rand(np.float64, n, n, 127, seed=1)
        B7 = g.rand(np.float64, n, n, 127, seed=1, offset=n * n)
        C7a = g.zeros(np.float64, n, n, 127)
        C7b = g.zeros(np.float64, n, n, 127)
        C7c = g.zeros(np.float64, n, n, 127)
        p127_limb_ms, _, _, _ = time_gemm(g, ctx, A7, B7, C7a, g.FORM_LIMB_I8, a.steps, a.warmup, "gemm_limb")
        # one int8 product modulo 127 itself (multi-modular form, one plane)
        p127_rns_ms, _, p127_rns_kms, _ = time_gemm(g, ctx, A7, B7, C7c, g.FORM_RNS_I8, a.steps, a.warmup,
                                                    "gemm_rns")
        p127_plain_ms, _, _, _ = time_gemm(g, ctx, A7, B7, C7b, g.FORM_PLAIN_F64, max(2, a.steps // 2), 1,
                                           "gemm_f64")
        c7b = g.Array(C7b)
        extra.update(p127_limb_gops=round(ops / (p127_limb_ms * 1e-3) / 1e9, 2),
                     p127_rns_gops=round(ops / (p127_rns_ms * 1e-3) / 1e9, 2),
                     p127_rns_core_ms=round(p127_rns_kms, 4) if p127_rns_kms else None,
                     p127_plain_gops=round(ops / (p127_plain_ms * 1e-3) / 1e9, 2),
                     p127_limb_vs_plain=round(p127_plain_ms / p127_limb_ms, 3),
                     p127_rns_vs_plain=round(p127_plain_ms / p127_rns_ms, 3),
                     p127_max_abs_delta_limb_vs_plain=float(np.abs(g.Array(C7a) - c7b).max()),
                     p127_max_abs_delta_rns_vs_plain=float(np.abs(g.Array(C7c) - c7b).max()))
        del A7, B7, C7a, C7b, C7c, c7b

        # 1024^2 plain GEMM, p = 65521, seed 0
        A1 = g.rand(np.float64, 1024, 1024, 65521, seed=0)
        B1 = g.rand(np.float64, 1024, 1024, 65521, seed=0, offset=1024 * 1024)
        C1 = g.zeros(np.float64, 1024, 1024, 65521)
        c1_ms, _, _, _ = time_gemm(g, ctx, A1, B1, C1, g.FORM_PLAIN_F64, 20, 3, "gemm_f64")
        extra.update(c1_plain_1024_gops=round(2.0 * 1024 ** 3 / (c1_ms * 1e-3) / 1e9, 2),
                     c1_plain_1024_ms=round(c1_ms, 4))
        del A1, B1, C1

        # matrix-vector product 8192 x 8192 mod 2^26-5: one HBM pass over A
        Av = g.rand(np.float64, n, n, p, seed=1)
        xv = g.rand(np.float64, n, 1, p, seed=1, offset=n * n)
        zv = g.zeros(np.float64, n, 1, p)
        for _ in range(3):
            g.mul_(zv, Av, xv)
        ctx.synchronize()
        ctx.event_record(4)
        for _ in range(10):
            g.mul_(zv, Av, xv)
        ctx.event_record(5)
        gemv_ms = ctx.event_elapsed(4, 5) / 10
        extra.update(gemv_8192_ms=round(gemv_ms, 4),
                     gemv_8192_hbm_gbs=round(8.0 * n * n / (gemv_ms * 1e-3) / 1e9, 1))
        del Av, xv, zv

        extra.update(bench_pluq(g, ctx))
        bg = BatchedGemm(g, ctx)
        extra.update(bg.run(reps=5))
        be = BatchedElim(g, ctx)
        extra.update(be.run())
    else:
        if a.calib_copy:
            ctx.check(ctx.lib.modp_memcpy_d2d(ctx.h, C_rns.ptr, A.ptr, 8 * n * n))
            ctx.synchronize()
        del A, B

    # roofline of the dominant kernel, the int8 MFMA core of the RNS form
    # (all moduli in one launch): algorithmic work 2 m n k per launch over
    # the mean launch time (HIP events around each launch on the library
    # stream)
    r_mod = 9  # moduli for p = 2^26 - 5 and k = 8192: 4 k floor(p/2)^2 < prod q_j
    achieved = ops / (rns_kms * 1e-3) / 1e12 if rns_kms else None     # TOP/s
    issued = r_mod * achieved if achieved else None                    # one int8 product per modulus
    traffic, traffic_src = None, None
    tfile = os.path.join(REPO, "profiles", "gemm_rns_8192_traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("bytes_per_launch")
        traffic_src = ("profiles/gemm_rns_8192_traffic.md: rocprofv3 FETCH_SIZE and WRITE_SIZE runs of the "
                       "same kernel, calibrated on a device copy (not measured in this run)")
    try:
        measured = {"int8_mfma_32x32x32_tops": round(ctx.probe_peak(3) / 1e12, 1),
                    "int8_mfma_16x16x64_tops": round(ctx.probe_peak(0) / 1e12, 1),
                    "fp64_mfma_tflops": round(ctx.probe_peak(1) / 1e12, 2),
                    "hbm_copy_gbs": round(ctx.probe_peak(2) / 1e9, 1)}
    except Exception as e:  # reported, never fatal
        measured = {"error": str(e)}
    m32 = measured.get("int8_mfma_32x32x32_tops")
    roof = {"bound": "mfma", "kernel": "k_rns_gemm2<256x256 tile, 2x4 waves> (9 moduli per launch)",
            "achieved": round(achieved, 2) if achieved else None,
            "peak": INT8_PEAK_TOPS, "unit": "TFLOP/s",
            "frac": round(achieved / INT8_PEAK_TOPS, 4) if achieved else None,
            "traffic": traffic, "traffic_source": traffic_src,
            # the core kernel reads r int8 residue planes of A and of B and
            # writes r residue planes of C once each
            "algorithmic_bytes": int(3 * r_mod * n * n),
            "measured_peaks": measured,
            "frac_of_measured": round(achieved / m32, 4) if achieved and m32 else None,
            "mfma_issue": {"int8_products": r_mod,
                           "issued_tops": round(issued, 1) if issued else None,
                           "issued_frac_of_peak": round(issued / INT8_PEAK_TOPS, 4) if issued else None,
                           "issued_frac_of_measured": round(issued / m32, 4) if issued and m32 else None},
            "note": "achieved = 2*8192^3 mod-p ops per k_rns_gemm2 launch / mean launch time (HIP events on the "
                    "library stream); TFLOP/s here counts int8 ops (TOP/s); frac_of_measured is against the "
                    "fastest int8 MFMA instruction measured on this device (v_mfma_i32_32x32x32_i8)"}

    cpu, parity = None, {}
    if not a.no_cpu:
        try:
            cpu, parity = cpu_baseline(g, n, p, 1, C_rns, bg, be)
        except Exception as e:  # the checker is optional at run time
            cpu = {"error": str(e)}

    res = {
        "metric": "mod-p matmul GOPS n=8192 int8-limb MFMA vs FP64 plain (bit-exact); PLUQ sec",
        "value": round(ops / (rns_ms * 1e-3) / 1e9, 2),
        "unit": "GOPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(rns_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic",
        "config": {"workload": f"mod-p GEMM {n}x{n}x{n}, p=2^26-5, residues uniform from seed 1, int8 MFMA "
                               "multi-modular limbs (residues modulo 9 pairwise coprime moduli <= 256), int32 "
                               "accumulation, FP64 Chinese-remainder recombination; each step includes the "
                               "limb split, the 9 int8 products and the recombination",
                   "n": n, "p": p, "form": "rns_i8", "parallelism": "single GPU"},
        "wall_ms_per_step": round(rns_wall, 4),
        "gemm_core_ms": round(rns_kms, 4) if rns_kms else None,
        **extra,
        "parity": parity,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
