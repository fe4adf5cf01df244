"""Builds the CPU model of the library: the unchanged csrc/ sources compiled
with g++ against tests/cpu_model/include/hip/hip_runtime.h (a CPU model of
the HIP device language and runtime), with AddressSanitizer and
UndefinedBehaviorSanitizer, into tests/cpu_model/_build/libsfntt_model.so.

    python tests/cpu_model/build_model.py [-j N]

Test infrastructure only: the product library is small-field-metal-ntt_amd/
sfntt/libsfntt.so, built by hipcc for gfx950.  A unit is recompiled when its
source, a header or this script is newer than its object.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "small-field-metal-ntt_amd", "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "libsfntt_model.so")
CXX = os.environ.get("CXX", "g++")

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]
FLAGS = ["-std=c++17", "-O1", "-g", "-fPIC", "-Wall", "-Wno-unknown-pragmas", "-pthread", *SAN,
         "-I", os.path.join(HERE, "include"), "-I", CSRC, "-I", os.path.join(ROOT, "include")]

UNITS = ([(os.path.join(CSRC, "kernels.hip"), [f"-DSFN_PART={k}"], f"kernels_p{k}.o") for k in range(8)] +
         [(os.path.join(CSRC, "fast.hip"), [f"-DSFN_PART={k}"], f"fast_p{k}.o") for k in range(8)] +
         [(os.path.join(CSRC, s), [], s.replace(".cpp", ".o")) for s in
          ("runtime.cpp", "tables.cpp", "batch.cpp", "server.cpp")] +
         [(os.path.join(HERE, "model_runtime.cpp"), [], "model_runtime.o")])


def _newest_dep() -> float:
    deps = glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(HERE, "include", "hip", "*.h"))
    deps += [os.path.join(ROOT, "include", "sfntt.h"), os.path.abspath(__file__)]
    return max(os.path.getmtime(d) for d in deps)


def _compile(unit, newest):
    src, extra, obj = unit
    out = os.path.join(BUILD, obj)
    if os.path.exists(out) and os.path.getmtime(out) >= max(newest, os.path.getmtime(src)):
        return out, 0.0
    cmd = [CXX, *FLAGS, *extra, "-x", "c++", "-c", src, "-o", out]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stderr[-4000:]}")
    return out, time.time() - t0


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    newest = _newest_dep()
    jobs = jobs or min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda u: _compile(u, newest), UNITS))
    if verbose:
        for o, t in results:
            if t:
                print(f"  {os.path.basename(o)}: {t:.1f}s")
    objs = [o for o, _ in results]
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < max(os.path.getmtime(o) for o in objs):
        tmp = LIB + ".tmp"
        r = subprocess.run([CXX, "-shared", "-pthread", *SAN, "-o", tmp, *objs], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed\n{r.stderr[-4000:]}")
        os.replace(tmp, LIB)
    return LIB


def asan_runtime() -> str:
    """Path of the compiler's ASan runtime (preloaded into the Python process
    that loads the model library)."""
    return subprocess.run([CXX, "-print-file-name=libasan.so"], capture_output=True, text=True,
                          check=True).stdout.strip()


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    t0 = time.time()
    path = build(a.j, verbose=True)
    print(f"built {path} in {time.time() - t0:.1f}s")
    sys.exit(0)
