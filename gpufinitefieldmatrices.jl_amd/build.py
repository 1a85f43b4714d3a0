"""Build libmodp_hip.so (HIP, gfx950) in-tree.

    python gpufinitefieldmatrices.jl_amd/build.py [--jobs N] [--only elim.hip ...]

--only recompiles the named sources and relinks with the objects already in
_build/ (a quick rebuild after editing one file that no other file includes).

Each csrc/*.hip file is compiled separately (in parallel) with hipcc for
gfx950 only, then linked into gpufinitefieldmatrices/libmodp_hip.so next to the
Python host package that loads it.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "gpufinitefieldmatrices")
BUILD_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libmodp_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-result",
    # packed FP32 instructions disabled for every device pass
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(BUILD_DIR, src.replace(".hip", ".o"))
    cmd = [HIPCC, *COMMON_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in (r.stderr or "").splitlines()
                    if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    if err.strip():
        sys.stderr.write(err + "\n")
    return obj


def build(jobs: int | None = None, only: list | None = None) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = _sources()
    todo = srcs if not only else [s for s in srcs if s in only]
    if only and len(todo) != len(only):
        raise RuntimeError(f"unknown sources in {only}")
    jobs = jobs or min(len(todo), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_compile, todo))
    objs = [os.path.join(BUILD_DIR, s.replace(".hip", ".o")) for s in srcs]
    missing = [o for o in objs if not os.path.exists(o)]
    if missing:
        raise RuntimeError(f"objects missing (run a full build): {missing}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--only", nargs="+", default=None)
    a = ap.parse_args()
    print(build(a.jobs, a.only))
