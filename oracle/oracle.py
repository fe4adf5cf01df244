"""ctypes front-end of the CPU oracle (liboracle.so, built from sfntt_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker (or the timed CPU
baseline), never as the product path.  Every wrapper restates a reference
CPU function; the citations sit on the C definitions in sfntt_oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

M31_P = 0x7FFFFFFF
BB_P = 0x78000001
KB_P = 0x7F000001
FIELD_M31, FIELD_BB, FIELD_KB = 0, 1, 2
LAYOUT_SOA, LAYOUT_ROW_MAJOR = 0, 1

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u32p = ctypes.POINTER(ctypes.c_uint32)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        u32, u64, sz, i = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_int
        for name, res, args in [
            ("orc_m31_reduce", u32, [u64]), ("orc_m31_add", u32, [u32, u32]),
            ("orc_m31_sub", u32, [u32, u32]), ("orc_m31_mul", u32, [u32, u32]),
            ("orc_m31_inv", u32, [u32]),
            ("orc_bb_monty_reduce", u32, [u64]), ("orc_bb_to_monty", u32, [u32]),
            ("orc_bb_from_monty", u32, [u32]), ("orc_bb_add", u32, [u32, u32]),
            ("orc_bb_sub", u32, [u32, u32]), ("orc_bb_mul", u32, [u32, u32]),
            ("orc_bb_reduce", u32, [u64]), ("orc_bb_inv", u32, [u32]),
            ("orc_bb_two_adic_generator", u32, [u32]),
            ("orc_kb_monty_reduce", u32, [u64]), ("orc_kb_to_monty", u32, [u32]),
            ("orc_kb_from_monty", u32, [u32]), ("orc_kb_add", u32, [u32, u32]),
            ("orc_kb_sub", u32, [u32, u32]), ("orc_kb_mul", u32, [u32, u32]),
            ("orc_kb_inv", u32, [u32]), ("orc_kb_two_adic_generator", u32, [u32]),
            ("orc_kb_twiddles", i, [u32, i, _u32p]),
            ("orc_kb_forward", i, [_u32p, sz]), ("orc_kb_inverse", i, [_u32p, sz]),
            ("orc_circle_odds_point", None, [u32, u64, _u32p, _u32p]),
            ("orc_bb_twiddles", i, [u32, i, _u32p]), ("orc_m31_twiddles", i, [u32, i, _u32p]),
            ("orc_m31_forward", i, [_u32p, sz]), ("orc_m31_inverse", i, [_u32p, sz]),
            ("orc_bb_forward", i, [_u32p, sz]), ("orc_bb_inverse", i, [_u32p, sz]),
            ("orc_ntt_batch", i, [i, i, u32, u32, i, _u32p, i]),
            ("orc_coset_lde_batch", i, [_u32p, u32, u32, u32, u32, i, _u32p, i]),
            ("orc_lcg_fill", u64, [u64, u32, i, _u32p, sz]),
            ("orc_splitmix64", u64, [ctypes.POINTER(ctypes.c_uint64)]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u32p)


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


# ---------------------------------------------------------------- scalars
def m31_mul(a, b): return lib().orc_m31_mul(a, b)
def m31_add(a, b): return lib().orc_m31_add(a, b)
def m31_sub(a, b): return lib().orc_m31_sub(a, b)
def m31_inv(a): return lib().orc_m31_inv(a)
def m31_reduce(v): return lib().orc_m31_reduce(v)
def bb_mul(a, b): return lib().orc_bb_mul(a, b)
def bb_add(a, b): return lib().orc_bb_add(a, b)
def bb_sub(a, b): return lib().orc_bb_sub(a, b)
def bb_inv(a): return lib().orc_bb_inv(a)
def bb_reduce(v): return lib().orc_bb_reduce(v)
def bb_to_monty(v): return lib().orc_bb_to_monty(v)
def bb_from_monty(v): return lib().orc_bb_from_monty(v)
def bb_monty_reduce(t): return lib().orc_bb_monty_reduce(t)
def bb_two_adic_generator(log_n): return lib().orc_bb_two_adic_generator(log_n)
def kb_mul(a, b): return lib().orc_kb_mul(a, b)
def kb_add(a, b): return lib().orc_kb_add(a, b)
def kb_sub(a, b): return lib().orc_kb_sub(a, b)
def kb_inv(a): return lib().orc_kb_inv(a)
def kb_to_monty(v): return lib().orc_kb_to_monty(v)
def kb_from_monty(v): return lib().orc_kb_from_monty(v)
def kb_monty_reduce(t): return lib().orc_kb_monty_reduce(t)
def kb_two_adic_generator(log_n): return lib().orc_kb_two_adic_generator(log_n)


def circle_odds_point(log_size: int, index: int):
    x, y = ctypes.c_uint32(), ctypes.c_uint32()
    lib().orc_circle_odds_point(log_size, index, ctypes.byref(x), ctypes.byref(y))
    return x.value, y.value


# --------------------------------------------------------------- tables
def bb_twiddles(log_n: int, inverse: bool = False) -> np.ndarray:
    out = np.zeros(max((1 << log_n) - 1, 1), dtype=np.uint32)
    assert lib().orc_bb_twiddles(log_n, int(inverse), _ptr(out)) == 0
    return out[: (1 << log_n) - 1]


def kb_twiddles(log_n: int, inverse: bool = False) -> np.ndarray:
    out = np.zeros(max((1 << log_n) - 1, 1), dtype=np.uint32)
    assert lib().orc_kb_twiddles(log_n, int(inverse), _ptr(out)) == 0
    return out[: (1 << log_n) - 1]


def m31_twiddles(log_n: int, inverse: bool = False) -> np.ndarray:
    out = np.zeros(max((1 << log_n) - 1, 1), dtype=np.uint32)
    assert lib().orc_m31_twiddles(log_n, int(inverse), _ptr(out)) == 0
    return out[: (1 << log_n) - 1]


# ----------------------------------------------------------- transforms
class InvalidSize(ValueError):
    """Mirrors NttError::InvalidSize (src/ntt/mod.rs:53)."""


def _single(fn, data) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(data, dtype=np.uint32)).copy()
    if fn(_ptr(a) if a.size else _u32p(), a.size) != 0:
        raise InvalidSize(a.size)
    return a


def m31_forward(data): return _single(lib().orc_m31_forward, data)
def m31_inverse(data): return _single(lib().orc_m31_inverse, data)
def bb_forward(data): return _single(lib().orc_bb_forward, data)
def bb_inverse(data): return _single(lib().orc_bb_inverse, data)
def kb_forward(data): return _single(lib().orc_kb_forward, data)
def kb_inverse(data): return _single(lib().orc_kb_inverse, data)


def ntt_batch(field: int, inverse: bool, log_n: int, batch: int, data, layout: int = LAYOUT_SOA,
              threads: int | None = None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(data, dtype=np.uint32)).copy()
    if a.size != (batch << log_n):
        raise InvalidSize(a.size)
    rc = lib().orc_ntt_batch(field, int(inverse), log_n, batch, layout, _ptr(a),
                             threads or default_threads())
    if rc != 0:
        raise InvalidSize(a.size)
    return a


def coset_lde_batch(data, log_n: int, batch: int, added_bits: int, shift_raw: int = 0,
                    layout: int = LAYOUT_SOA, threads: int | None = None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(data, dtype=np.uint32))
    if a.size != (batch << log_n):
        raise InvalidSize(a.size)
    out = np.empty(batch << (log_n + added_bits), dtype=np.uint32)
    rc = lib().orc_coset_lde_batch(_ptr(a), log_n, batch, added_bits, shift_raw, layout,
                                   _ptr(out), threads or default_threads())
    if rc != 0:
        raise InvalidSize(a.size)
    return out


# ------------------------------------------------------------ generators
def lcg_fill(seed: int, n: int, field: int) -> np.ndarray:
    """Reference LCG data: M31 canonical values, BabyBear or KoalaBear Montgomery raw."""
    out = np.empty(n, dtype=np.uint32)
    p = {FIELD_M31: M31_P, FIELD_BB: BB_P, FIELD_KB: KB_P}[field]
    lib().orc_lcg_fill(seed, p, field, _ptr(out) if n else _u32p(), n)
    return out


def splitmix64_stream(seed: int):
    st = ctypes.c_uint64(seed)
    while True:
        yield lib().orc_splitmix64(ctypes.byref(st))
