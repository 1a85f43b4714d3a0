"""Benchmark: mod-p GEMM at n = 8192, p = 2^26 - 5, int8-limb MFMA form vs the
plain FP64-MFMA form (bit-exact), plus PLUQ seconds at 4096^2 mod 65521.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

One "step" is one C = A*B mod p over 8192^3 residues resident in HBM (the
limb split, the int8 MFMA products and the recombination all inside the
step).  Prints one JSON line.
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
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-pluq", action="store_true")
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
    """The oracle (C, OpenMP over the host cores) on a bounded sample of the
    same workload: a 512-row slice of A times the full B (512 x n x n)."""
    import oracle as O
    rows = 512
    A = O.draw_matrix(rows, n, p, seed=seed)
    B = O.draw_matrix(n, n, p, seed=seed, offset=n * n)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    O.gemm(A, B, p)
    dt = time.perf_counter() - t0
    return dict(value=round(2.0 * rows * n * n / dt / 1e9, 3), unit="GOPS", cores=cores, kind="port",
                sample=f"oracle_gemm {rows}x{n}x{n} (row slice of the {n}^3 product), p={p}, {dt:.2f} s")


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import gpufinitefieldmatrices as g
    ctx = g.context()
    n = a.n
    p = 2 ** 26 - 5
    ops = 2.0 * n * n * n

    A = g.rand(np.float64, n, n, p, seed=1)
    B = g.rand(np.float64, n, n, p, seed=1, offset=n * n)
    C_limb = g.zeros(np.float64, n, n, p)
    C_plain = g.zeros(np.float64, n, n, p)

    limb_ms, limb_wall, limb_kms, _ = time_gemm(g, ctx, A, B, C_limb, g.FORM_LIMB_I8, a.steps, a.warmup, "gemm_limb")
    plain_ms, _, plain_kms, _ = time_gemm(g, ctx, A, B, C_plain, g.FORM_PLAIN_F64, max(2, a.steps // 2),
                                          1, "gemm_f64")
    maxdelta = float(np.abs(g.Array(C_limb) - g.Array(C_plain)).max())

    # single-limb prime p = 127
    A7 = g.rand(np.float64, n, n, 127, seed=1)
    B7 = g.rand(np.float64, n, n, 127, seed=1, offset=n * n)
    C7a = g.zeros(np.float64, n, n, 127)
    C7b = g.zeros(np.float64, n, n, 127)
    p127_limb_ms, _, p127_kms, _ = time_gemm(g, ctx, A7, B7, C7a, g.FORM_LIMB_I8, a.steps, a.warmup, "gemm_limb")
    p127_plain_ms, _, _, _ = time_gemm(g, ctx, A7, B7, C7b, g.FORM_PLAIN_F64, max(2, a.steps // 2), 1, "gemm_f64")
    maxdelta127 = float(np.abs(g.Array(C7a) - g.Array(C7b)).max())
    del A7, B7, C7a, C7b

    out_pluq = {}
    if not a.no_pluq:
        q = 65521
        nq = 4096
        Aq = g.rand(np.float64, nq, nq, q, seed=2)
        for name, form in (("limb", g.FORM_LIMB_I8), ("plain", g.FORM_PLAIN_F64)):
            g.pluq(Aq, form=form)  # warm-up (workspace allocation)
            ctx.synchronize()
            ts = []
            for _ in range(2):
                t0 = time.perf_counter()
                _, _, _, _, rk = g.pluq(Aq, form=form)
                ctx.synchronize()
                ts.append(time.perf_counter() - t0)
            out_pluq[f"pluq_sec_{name}"] = round(min(ts), 5)
            out_pluq["pluq_rank"] = rk

    # roofline of the dominant kernel: the int8 limb MFMA GEMM core.
    # issued int8 MFMA ops per launch = (limb pairs) x 2 m n k (padded k)
    L = 4
    issued = L * L * 2.0 * n * n * n
    achieved_tops = issued / (limb_kms * 1e-3) / 1e12 if limb_kms else None
    traffic = None
    tfile = os.path.join(REPO, "profiles", "gemm_limb_8192_traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("bytes_per_launch")
    roof = {"bound": "mfma", "kernel": "k_gemm_limb<4,64,64>", "achieved": round(achieved_tops, 2) if achieved_tops else None,
            "peak": INT8_PEAK_TOPS, "unit": "TFLOP/s", "frac": round(achieved_tops / INT8_PEAK_TOPS, 4) if achieved_tops else None,
            "traffic": traffic,
            "note": "achieved = 16 limb pairs x 2*8192^3 int8 ops per launch / mean launch time (HIP events)"}

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
                               "4x7-bit int8 limbs, int32 MFMA accumulation, FP64 Barrett recombine",
                   "n": n, "p": p, "form": "limb_i8", "parallelism": "single GPU"},
        "plain_f64_gops": round(ops / (plain_ms * 1e-3) / 1e9, 2),
        "limb_vs_plain": round(plain_ms / limb_ms, 3),
        "max_abs_delta": maxdelta,
        "p127_limb_gops": round(ops / (p127_limb_ms * 1e-3) / 1e9, 2),
        "p127_plain_gops": round(ops / (p127_plain_ms * 1e-3) / 1e9, 2),
        "p127_limb_vs_plain": round(p127_plain_ms / p127_limb_ms, 3),
        "p127_max_abs_delta": maxdelta127,
        "wall_ms_per_step": round(limb_wall, 4),
        "gemm_core_ms": round(limb_kms, 4) if limb_kms else None,
        "plain_core_ms": round(plain_kms, 4) if plain_kms else None,
        **out_pluq,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
