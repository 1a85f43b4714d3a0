"""Benchmark: mod-p GEMM at n = 8192, p = 2^26 - 5, int8-limb MFMA form vs the
plain FP64-MFMA form (bit-exact), plus PLUQ seconds at 4096^2 mod 65521 and
the other workloads of the hot path (1024^2 plain GEMM, batched products,
batched eliminations).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--only gemm]

One "step" is one C = A*B mod p over 8192^3 residues resident in HBM (the
limb split, the int8 MFMA products and the recombination all inside the
step).  Prints one JSON line.  Inputs are generated on the device by the
library's seeded generator.
"""
from __future__ import annotations

import argparse
import json
import os
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


def cpu_baseline(n, p, seed):
    """The oracle (C, OpenMP over the host cores) on the same workload: the
    full n^3 product of the same seeded residues."""
    import oracle as O
    A = O.draw_matrix(n, n, p, seed=seed)
    B = O.draw_matrix(n, n, p, seed=seed, offset=n * n)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    O.gemm(A, B, p)
    dt = time.perf_counter() - t0
    return dict(value=round(2.0 * n ** 3 / dt / 1e9, 3), unit="GOPS", cores=cores, kind="port",
                sample=f"oracle_gemm full {n}x{n}x{n}, p={p}, seed {seed}: {dt:.2f} s")


def largest_primes_below(bound, count):
    def is_prime(x):
        if x < 2:
            return False
        for d in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
            if x % d == 0:
                return x == d
        dd, s = x - 1, 0
        while dd % 2 == 0:
            dd //= 2
            s += 1
        for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
            y = pow(a, dd, x)
            if y in (1, x - 1):
                continue
            for _ in range(s - 1):
                y = y * y % x
                if y == x - 1:
                    break
            else:
                return False
        return True
    out, x = [], bound - 1
    while len(out) < count:
        if is_prime(x):
            out.append(x)
        x -= 1
    return out


def bench_batched_gemm(g, ctx, reps):
    """10,000 independent products, m,n,k ~ U[16,256] (seed 3), each under its
    own prime from the 64 largest primes below 2^23."""
    rng = np.random.default_rng(3)
    primes = largest_primes_below(2 ** 23, 64)
    count = 10000
    shp = rng.integers(16, 257, size=(count, 3))
    pr = [primes[i] for i in rng.integers(0, 64, size=count)]
    elems = int(sum(m * k + k * n + m * n for m, n, k in shp))
    buf = ctx.malloc(8 * elems)
    jobs = (g.GemmJob * count)()
    off = 0
    ops = 0
    for i, ((m, n, k), p) in enumerate(zip(shp.tolist(), pr)):
        a, b, c = buf.ptr + 8 * off, buf.ptr + 8 * (off + m * k), buf.ptr + 8 * (off + m * k + k * n)
        off += m * k + k * n + m * n
        ctx.check(ctx.lib.modp_rand(ctx.h, 0, m, k, a, m, p, (3 << 32) + 2 * i, 0))
        ctx.check(ctx.lib.modp_rand(ctx.h, 0, k, n, b, k, p, (3 << 32) + 2 * i + 1, 0))
        jobs[i] = g.GemmJob(m, n, k, p, a, m, b, k, c, m)
        ops += 2 * m * n * k
    import ctypes
    h = ctypes.c_void_p()
    # the plan's tile policy is autotuned (candidate plans built, run and
    # timed before the timed region); the tuner's cache is left as it was
    ctx.autotune_mode(g.TUNE_ON_MISS)
    try:
        ctx.check(ctx.lib.modp_gemm_batched_plan(ctx.h, 0, jobs, count, ctypes.byref(h)))
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
    del buf
    return dict(batched_gemm_gops=round(ops / (ms * 1e-3) / 1e9, 2), batched_gemm_ms=round(ms, 4),
                batched_gemm_hbm_gbs=round(8.0 * elems / (ms * 1e-3) / 1e9, 1))


def bench_batched_elim(g, ctx):
    """4096 independent pluq problems, rows/cols ~ U[8,128] (seed 4), prime
    per instance from {2,3,7,251,65521}; 25% rank-deficient (a random row
    subset zeroed, then left-multiplied by an invertible matrix)."""
    import ctypes
    rng = np.random.default_rng(4)
    primes = [2, 3, 7, 251, 65521]
    count = 4096
    inst = []
    for i in range(count):
        m, n = (int(x) for x in rng.integers(8, 129, size=2))
        p = primes[int(rng.integers(0, 5))]
        A = rng.integers(0, p, size=(m, n))
        if i % 4 == 0:
            A[rng.choice(m, size=int(rng.integers(1, m)), replace=False), :] = 0
            Ad = g.CuModMatrix(A, p, elem_type=np.float64)
            while True:
                S = g.CuModMatrix(rng.integers(0, p, size=(m, m)), p, elem_type=np.float64)
                if g.pluq(S)[4] == m:
                    break
            A = g.Array(S * Ad).astype(np.int64)
        inst.append((A, p))
    # one device arena for inputs and outputs
    sizes = [(A.shape[0], A.shape[1]) for A, _ in inst]
    nA = sum(m * n for m, n in sizes)
    nL = sum(m * m for m, _ in sizes)
    arena = ctx.malloc(8 * (2 * nA + nL))
    info = ctx.malloc(8 * sum(2 * min(m, n) + 4 * n + 3 for m, n in sizes))
    host = np.zeros(nA, dtype=np.float64)
    jobs = (g.ElimJob * count)()
    oa = ou = ol = oi = 0
    for i, ((A, p), (m, n)) in enumerate(zip(inst, sizes)):
        host[oa:oa + m * n] = A.astype(np.float64).ravel(order="F")
        a = arena.ptr + 8 * oa
        u = arena.ptr + 8 * (nA + ou)
        l = arena.ptr + 8 * (2 * nA + ol)
        ri = info.ptr + 8 * oi
        ci = ri + 8 * 2 * min(m, n)
        cn = ci + 8 * 4 * n
        jobs[i] = g.ElimJob(m, n, p, a, m, u, m, l, m, ri, ci, cn)
        oa += m * n
        ou += m * n
        ol += m * m
        oi += 2 * min(m, n) + 4 * n + 3
    ctx.check(ctx.lib.modp_memcpy_h2d(ctx.h, arena.ptr, host.ctypes.data, host.nbytes))
    ctx.check(ctx.lib.modp_elim_batched(ctx.h, 0, 0, jobs, count))
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        ctx.check(ctx.lib.modp_elim_batched(ctx.h, 0, 0, jobs, count))
        ts.append(time.perf_counter() - t0)
    ops = sum(2 * min(m, n) * m * n for m, n in sizes)
    ms = min(ts) * 1e3
    return dict(batched_pluq_ms=round(ms, 3), batched_pluq_instances=count,
                batched_pluq_hbm_gbs=round(8.0 * (2 * nA + nL) / (ms * 1e-3) / 1e9, 1),
                batched_pluq_gops_upper=round(ops / (ms * 1e-3) / 1e9, 2))


def bench_pluq(g, ctx):
    """PLUQ of 4096^2 mod 65521 (seed 2) with both forms of the trailing
    update, and the rank-3000 product (4096x3000)(3000x4096) by PLUQ and RREF."""
    q, nq = 65521, 4096
    extra = {}
    Aq = g.rand(np.float64, nq, nq, q, seed=2)
    for name, form in (("limb", g.FORM_LIMB_I8), ("plain", g.FORM_PLAIN_F64)):
        g.pluq(Aq, form=form)
        ctx.synchronize()
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            _, _, _, _, rk = g.pluq(Aq, form=form)
            ctx.synchronize()
            ts.append(time.perf_counter() - t0)
        extra[f"pluq_sec_{name}"] = round(min(ts), 5)
        extra["pluq_rank"] = rk
    extra["pluq_limb_vs_plain"] = round(extra["pluq_sec_plain"] / extra["pluq_sec_limb"], 3)
    # rank-3000 product (4096x3000)(3000x4096)
    X = g.rand(np.float64, nq, 3000, q, seed=2)
    Y = g.rand(np.float64, 3000, nq, q, seed=2, offset=nq * 3000)
    A3 = X * Y
    del X, Y
    g.pluq(A3, form=g.FORM_LIMB_I8)
    t0 = time.perf_counter()
    _, _, _, _, rk3 = g.pluq(A3, form=g.FORM_LIMB_I8)
    ctx.synchronize()
    extra.update(pluq_rank3000_sec=round(time.perf_counter() - t0, 5), pluq_rank3000_rank=rk3)
    t0 = time.perf_counter()
    _, _, rk3r = g.rref_full(A3, form=g.FORM_LIMB_I8)
    ctx.synchronize()
    extra.update(rref_rank3000_sec=round(time.perf_counter() - t0, 5))
    del A3
    # A^{-1} of the full-rank 4096^2 matrix (PLUQ, two triangular inverses,
    # permutations, one product)
    g.inverse(Aq)
    ctx.synchronize()
    t0 = time.perf_counter()
    g.inverse(Aq)
    ctx.synchronize()
    extra.update(inverse_4096_sec=round(time.perf_counter() - t0, 5))
    del Aq

    return extra


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
        out = bench_pluq(g, ctx) if a.only == "pluq" else {**bench_batched_gemm(g, ctx, reps=5),
                                                            **bench_batched_elim(g, ctx)}
        print(json.dumps({"section": a.only, **out}))
        return
    n = a.n
    p = 2 ** 26 - 5
    ops = 2.0 * n * n * n

    A = g.rand(np.float64, n, n, p, seed=1)
    B = g.rand(np.float64, n, n, p, seed=1, offset=n * n)
    C_limb = g.zeros(np.float64, n, n, p)
    limb_ms, limb_wall, limb_kms, _ = time_gemm(g, ctx, A, B, C_limb, g.FORM_LIMB_I8, a.steps, a.warmup, "gemm_limb")
    extra = {}
    if a.only == "all":
        C_plain = g.zeros(np.float64, n, n, p)
        plain_ms, _, plain_kms, _ = time_gemm(g, ctx, A, B, C_plain, g.FORM_PLAIN_F64, max(2, a.steps // 2), 1,
                                              "gemm_f64")
        maxdelta = float(np.abs(g.Array(C_limb) - g.Array(C_plain)).max())
        del C_plain
        extra.update(plain_f64_gops=round(ops / (plain_ms * 1e-3) / 1e9, 2),
                     limb_vs_plain=round(plain_ms / limb_ms, 3), max_abs_delta=maxdelta,
                     plain_core_ms=round(plain_kms, 4) if plain_kms else None,
                     plain_fp64_mfma_frac=round(ops / (plain_kms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4)
                     if plain_kms else None)
        del A, B
        # single-limb prime p = 127
        A7 = g.rand(np.float64, n, n, 127, seed=1)
        B7 = g.rand(np.float64, n, n, 127, seed=1, offset=n * n)
        C7a = g.zeros(np.float64, n, n, 127)
        C7b = g.zeros(np.float64, n, n, 127)
        p127_limb_ms, _, _, _ = time_gemm(g, ctx, A7, B7, C7a, g.FORM_LIMB_I8, a.steps, a.warmup, "gemm_limb")
        p127_plain_ms, _, _, _ = time_gemm(g, ctx, A7, B7, C7b, g.FORM_PLAIN_F64, max(2, a.steps // 2), 1,
                                           "gemm_f64")
        extra.update(p127_limb_gops=round(ops / (p127_limb_ms * 1e-3) / 1e9, 2),
                     p127_plain_gops=round(ops / (p127_plain_ms * 1e-3) / 1e9, 2),
                     p127_limb_vs_plain=round(p127_plain_ms / p127_limb_ms, 3),
                     p127_max_abs_delta=float(np.abs(g.Array(C7a) - g.Array(C7b)).max()))
        del A7, B7, C7a, C7b

        # 1024^2 plain GEMM, p = 65521, seed 0
        A1 = g.rand(np.float64, 1024, 1024, 65521, seed=0)
        B1 = g.rand(np.float64, 1024, 1024, 65521, seed=0, offset=1024 * 1024)
        C1 = g.zeros(np.float64, 1024, 1024, 65521)
        c1_ms, _, _, _ = time_gemm(g, ctx, A1, B1, C1, g.FORM_PLAIN_F64, 20, 3, "gemm_f64")
        extra.update(c1_plain_1024_gops=round(2.0 * 1024 ** 3 / (c1_ms * 1e-3) / 1e9, 2), c1_plain_1024_ms=round(c1_ms, 4))
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
        extra.update(bench_batched_gemm(g, ctx, reps=5))
        extra.update(bench_batched_elim(g, ctx))
    else:
        if a.calib_copy:
            ctx.check(ctx.lib.modp_memcpy_d2d(ctx.h, C_limb.ptr, A.ptr, 8 * n * n))
            ctx.synchronize()
        del A, B

    # roofline of the dominant kernel: the int8 limb MFMA GEMM core.
    # issued int8 MFMA ops per launch = (limb pairs) x 2 m n k
    L = 4
    issued = L * L * 2.0 * n * n * n
    achieved_tops = issued / (limb_kms * 1e-3) / 1e12 if limb_kms else None
    traffic = None
    tfile = os.path.join(REPO, "profiles", "gemm_limb_8192_traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("bytes_per_launch")
    # peaks measured on this device by the library's microbenchmarks
    try:
        measured = {"int8_mfma_tops": round(ctx.probe_peak(0) / 1e12, 1),
                    "fp64_mfma_tflops": round(ctx.probe_peak(1) / 1e12, 2),
                    "hbm_copy_gbs": round(ctx.probe_peak(2) / 1e9, 1)}
    except Exception as e:  # reported, never fatal
        measured = {"error": str(e)}
    roof = {"bound": "mfma", "kernel": "k_gemm_limb<4 digits, 64x64 tile>",
            "achieved": round(achieved_tops, 2) if achieved_tops else None,
            "peak": INT8_PEAK_TOPS, "unit": "TFLOP/s",
            "frac": round(achieved_tops / INT8_PEAK_TOPS, 4) if achieved_tops else None,
            "measured_peaks": measured,
            "frac_of_measured": (round(achieved_tops / measured["int8_mfma_tops"], 4)
                                 if achieved_tops and measured.get("int8_mfma_tops") else None),
            "traffic": traffic,
            "note": "achieved = 16 digit pairs x 2*8192^3 int8 ops per launch / mean launch time (HIP events); "
                    "TFLOP/s here counts int8 ops (TOP/s)"}

    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline(n, p, seed=1)
        except Exception as e:  # the checker is optional at run time
            cpu = {"error": str(e)}

    res = {
        "metric": "mod-p matmul GOPS n=8192 int8-limb MFMA vs FP64 plain (bit-exact); PLUQ sec",
        "value": round(ops / (limb_ms * 1e-3) / 1e9, 2),
        "unit": "GOPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(limb_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic",
        "config": {"workload": f"mod-p GEMM {n}x{n}x{n}, p=2^26-5, residues uniform from seed 1, "
                               "4 signed 8-bit limbs (base-256 digits) in int8, int32 MFMA accumulation, "
                               "FP64 Barrett recombine",
                   "n": n, "p": p, "form": "limb_i8", "parallelism": "single GPU"},
        "wall_ms_per_step": round(limb_wall, 4),
        "gemm_core_ms": round(limb_kms, 4) if limb_kms else None,
        **extra,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
