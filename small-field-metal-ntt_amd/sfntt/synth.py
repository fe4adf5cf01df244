"""Synthetic inputs: the reference's LCG, vectorised with numpy.

benches/ntt_benchmark.rs:53-61 (and :482-490, :653-661, src/lde/mod.rs:194-202):
    s = s * 6364136223846793005 + 1  (mod 2^64);  v = (s >> 33) % p
BabyBear inputs are stored as to_monty(v) = v * 2^32 mod p.
The k-th state is a^(k+1) * s0 + (1 + a + ... + a^k), so a chunk of states
is two wrapping uint64 products away from the chunk's starting state.
"""
from __future__ import annotations

import numpy as np

A = np.uint64(6364136223846793005)
M31_P = 0x7FFFFFFF
BB_P = 0x78000001
KB_P = 0x7F000001
_CHUNK = 1 << 20
_tables = None


def _chunk_tables():
    global _tables
    if _tables is None:
        with np.errstate(over="ignore"):
            pw = np.full(_CHUNK, A, dtype=np.uint64)
            pw = np.cumprod(pw, dtype=np.uint64)                 # a^(j+1)
            geo = np.concatenate([np.ones(1, np.uint64), pw[:-1]])  # a^j
            cs = np.cumsum(geo, dtype=np.uint64)                # sum_{i<=j} a^i
        _tables = (pw, cs)
    return _tables


def lcg_u32(seed: int, n: int, modulus: int, monty: bool, out: np.ndarray | None = None) -> np.ndarray:
    """n values of the reference LCG reduced mod `modulus`; Montgomery form if `monty`."""
    pw, cs = _chunk_tables()
    if out is None:
        out = np.empty(n, dtype=np.uint32)
    s = np.uint64(seed % (1 << 64))
    p = np.uint64(modulus)
    with np.errstate(over="ignore"):
        for off in range(0, n, _CHUNK):
            m = min(_CHUNK, n - off)
            st = pw[:m] * s + cs[:m]
            v = (st >> np.uint64(33)) % p
            if monty:
                v = (v << np.uint64(32)) % p
            out[off:off + m] = v.astype(np.uint32)
            s = st[m - 1]
    return out


def field_data(field: int, seed: int, n: int, out: np.ndarray | None = None) -> np.ndarray:
    """field 0 = M31 (canonical), 1 = BabyBear, 2 = KoalaBear (Montgomery raw)."""
    return lcg_u32(seed, n, (M31_P, BB_P, KB_P)[field], field != 0, out)


def splitmix64(state: int):
    """SplitMix64 (Steele, Lea, Flood 2014): returns (next_state, output)."""
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


def uneven_columns(seed: int = 0x5EED0004, count: int = 512, lo: int = 10, hi: int = 22):
    """Uneven batch draw: field ~ Bernoulli(1/2), log_n ~ U{lo..hi}, per-column LCG seed."""
    st = seed
    out = []
    for _ in range(count):
        st, a = splitmix64(st)
        st, b = splitmix64(st)
        st, c = splitmix64(st)
        out.append((a & 1, lo + b % (hi - lo + 1), c))
    return out
