"""Builds tests/cpp/_build/host_api_test (C++ host mirror test) with g++.

    python tests/cpp/build_host_test.py

Links the engine (small-field-metal-ntt_amd/sfntt/libsfntt.so) and the CPU
oracle (oracle/liboracle.so) through $ORIGIN-relative run paths, so the
binary runs wherever the tree is copied.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "host_api_test")
SRC = os.path.join(HERE, "host_api_test.cpp")
LIBDIR = os.path.join(ROOT, "small-field-metal-ntt_amd", "sfntt")
ORACLE = os.path.join(ROOT, "oracle")
HEADERS = [os.path.join(ROOT, "include", "sfntt.h"),
           os.path.join(ROOT, "small-field-metal-ntt_amd", "host", "sfntt.hpp")]


def stale() -> bool:
    if not os.path.exists(OUT):
        return True
    t = os.path.getmtime(OUT)
    return any(os.path.getmtime(f) > t for f in [SRC, *HEADERS])


def build(force: bool = True) -> str:
    if not force and not stale():
        return OUT
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall",
           "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "small-field-metal-ntt_amd", "host"),
           "-I", ORACLE, SRC, "-o", OUT,
           os.path.join(LIBDIR, "libsfntt.so"), os.path.join(ORACLE, "liboracle.so"),
           "-Wl,-rpath,$ORIGIN/../../../small-field-metal-ntt_amd/sfntt",
           "-Wl,-rpath,$ORIGIN/../../../oracle",
           "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib", "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host test build failed: {' '.join(cmd)}\n{r.stderr}")
    return OUT


if __name__ == "__main__":
    print(build())
    sys.exit(0)
