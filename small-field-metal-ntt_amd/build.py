"""Builds libsfntt.so (HIP kernels + C ABI runtime) for gfx950, in-tree.

    python small-field-metal-ntt_amd/build.py [-j N]

Objects go to small-field-metal-ntt_amd/_build/, the library to
small-field-metal-ntt_amd/sfntt/libsfntt.so.  Every call recompiles.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "sfntt", "libsfntt.so")
LOADGEN = os.path.join(HERE, "sfntt", "serve_load")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    # packed FP32 instructions are kept out of every device build
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", CSRC, "-I", os.path.join(os.path.dirname(HERE), "include"),
]

# (source, extra flags, object name)
UNITS = [("kernels.hip", [f"-DSFN_PART={k}"], f"kernels_p{k}.o") for k in range(8)] + [
    ("fast.hip", [f"-DSFN_PART={k}"], f"fast_p{k}.o") for k in range(8)] + [
    ("runtime.cpp", ["-x", "hip"], "runtime.o"),
    ("tables.cpp", ["-x", "hip"], "tables.o"),
    ("batch.cpp", ["-x", "hip"], "batch.o"),
    ("server.cpp", ["-x", "hip"], "server.o"),
]


def _compile(unit):
    src, extra, obj = unit
    path = os.path.join(CSRC, src)
    if not os.path.exists(path):
        return obj, None
    out = os.path.join(BUILD, obj)
    cmd = [HIPCC, *COMMON, *extra, "-c", path, "-o", out]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stderr}")
    return out, time.time() - t0


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, UNITS))
    objs = [o for o, t in results if t is not None]
    if verbose:
        for o, t in results:
            if t is not None:
                print(f"  {os.path.basename(o)}: {t:.1f}s")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the serving benchmark's native load generator (links the library)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-pthread", "-I", os.path.join(os.path.dirname(HERE), "include"),
           os.path.join(HERE, "serve_load.cpp"), "-L", os.path.dirname(LIB), "-lsfntt",
           "-Wl,-rpath,$ORIGIN", "-o", LOADGEN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"load generator build failed: {' '.join(cmd)}\n{r.stderr}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    t0 = time.time()
    path = build(a.j, verbose=True)
    print(f"built {path} in {time.time() - t0:.1f}s")
    sys.exit(0)
