"""Writes the golden fixtures of the GPU parity tests from the CPU oracle.

    python tests/golden/make_golden.py        (about a minute on 8 cores)

fixtures.npz   small input/output vectors: BabyBear and M31 forward/inverse
               at 2^4 and 2^9, coset LDEs (8,4,1) and (8,4,2) at the
               reference shift and at shift 31 (src/lde/mod.rs:194-271 cases)
digests.json   SHA-256 of the full outputs of the benchmark workloads, over
               the little-endian u32 bytes in the stated layout:
                 bb_fwd_2e16     BabyBear forward 2^16, seed 12345 + 2^16
                 m31_fwd_2e20    M31 forward 2^20, seed 12345 + 2^20
                 lde_2e20x256_*  coset LDE 2^20 x 256, 2x, seed 42 + 20 (inputs
                                 generated column by column): row-major with
                                 shift 31, and SoA with the reference shift;
                                 _natural: row-major, shift 31, natural row
                                 order on both sides
                 uneven_512_columns_forward
                                 512 uneven columns (SplitMix64 seed 0x5EED0004:
                                 field = draw1 & 1, log_n = 10 + draw2 % 13, LCG
                                 seed = draw3), forward, outputs concatenated
                 serving_slice_<k>
                                 job k of a 20-job slice of the serving stream at
                                 its stated sizes (serving_slice() below): a 2^20 x
                                 256 LDE at 4x and one at 2x, an M31 2^22 NTT, and
                                 17 jobs drawn from the stream's distribution
Inputs come from the oracle's restatement of the reference LCG; every output is
the oracle's.  The files are data only (numpy arrays without pickling, JSON).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as orc  # noqa: E402

M31, BB = orc.FIELD_M31, orc.FIELD_BB


def bitrev_index(bits: int) -> np.ndarray:
    """i -> i with its low `bits` bits reversed, for i < 2^bits"""
    i = np.arange(1 << bits, dtype=np.uint64)
    r = np.zeros_like(i)
    for b in range(bits):
        r |= ((i >> np.uint64(b)) & np.uint64(1)) << np.uint64(bits - 1 - b)
    return r.astype(np.int64)


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u4").tobytes()).hexdigest()


def uneven_shapes(seed=0x5EED0004, count=512):
    it = orc.splitmix64_stream(seed)
    out = []
    for _ in range(count):
        a, b, c = next(it), next(it), next(it)
        out.append((int(a & 1), 10 + int(b % 13), int(c)))
    return out


def serving_slice():
    """(kind, log_n, batch, added_bits, seed) of 20 jobs of the serving stream
    (60% BabyBear coset LDE 2^14..2^20 x 8..256 columns at 2x/4x with the
    reference shift, 40% M31 forward NTT 2^12..2^22): the three largest shapes,
    then 17 draws from the same distribution with LDE jobs capped at 2^17 x 64
    so the slice's outputs stay near 7 GiB.  Job k's input is the reference LCG
    with seed 0x5EED0005 + k (BabyBear Montgomery form for an LDE, M31 for an NTT)."""
    rng = np.random.default_rng(0x5EED0005)
    jobs = [("lde", 20, 256, 2), ("lde", 20, 256, 1), ("ntt", 22, 1, 0)]
    while len(jobs) < 20:
        if rng.random() < 0.6:
            jobs.append(("lde", int(rng.integers(14, 18)), int(rng.integers(8, 65)), int(rng.integers(1, 3))))
        else:
            jobs.append(("ntt", int(rng.integers(12, 23)), 1, 0))
    return [(k, n, b, a, 0x5EED0005 + i) for i, (k, n, b, a) in enumerate(jobs)]


def serving_slice_input(kind, log_n, batch, seed):
    return orc.lcg_fill(seed, batch << log_n, BB if kind == "lde" else M31)


def serving_slice_output(kind, log_n, batch, added_bits, x):
    if kind == "lde":
        return orc.coset_lde_batch(x, log_n, batch, added_bits)
    return orc.m31_forward(x)


def small_fixtures() -> dict:
    f = {}
    for field, name in ((BB, "bb"), (M31, "m31")):
        for log_n in (4, 9):
            x = orc.lcg_fill(777 + log_n, 1 << log_n, field)
            f[f"{name}_fwd_{log_n}_in"] = x
            f[f"{name}_fwd_{log_n}_out"] = orc.m31_forward(x) if field == M31 else orc.bb_forward(x)
            f[f"{name}_inv_{log_n}_out"] = orc.m31_inverse(x) if field == M31 else orc.bb_inverse(x)
    for log_n, batch, a in ((8, 4, 1), (8, 4, 2)):
        x = orc.lcg_fill(42, batch << log_n, BB)
        f[f"lde_{log_n}_{batch}_{a}_in"] = x
        f[f"lde_{log_n}_{batch}_{a}_ref_out"] = orc.coset_lde_batch(x, log_n, batch, a)
        f[f"lde_{log_n}_{batch}_{a}_g31_out"] = orc.coset_lde_batch(x, log_n, batch, a, orc.bb_to_monty(31))
    return f


def digests() -> dict:
    d = {}
    x = orc.lcg_fill(42, 8 << 12, BB)
    d["lde_12_8_1_ref"] = sha(orc.coset_lde_batch(x, 12, 8, 1))
    d["lde_12_8_1_g31"] = sha(orc.coset_lde_batch(x, 12, 8, 1, orc.bb_to_monty(31)))
    n = 1 << 16
    d["bb_fwd_2e16"] = sha(orc.bb_forward(orc.lcg_fill(12345 + n, n, BB)))
    n = 1 << 20
    d["m31_fwd_2e20"] = sha(orc.m31_forward(orc.lcg_fill(12345 + n, n, M31)))
    log_n, B, a = 20, 256, 1
    x = orc.lcg_fill(42 + log_n, B << log_n, BB)  # column by column (SoA order)
    y = orc.coset_lde_batch(x, log_n, B, a, 0)
    d["lde_2e20x256_2x_soa_refshift"] = sha(y)
    y = orc.coset_lde_batch(x, log_n, B, a, orc.bb_to_monty(31))  # SoA result, shift 31
    d["lde_2e20x256_2x_rowmajor_shift31"] = sha(y.reshape(B, 2 << log_n).T)
    # natural row order on both sides (input row i = P(g_n^i), output row k =
    # P(s g_N^k)): the same columns read as natural-order rows; the natural LDE
    # is the bit-reversed-order LDE between two row bit-reversals
    rn, rN = bitrev_index(log_n), bitrev_index(log_n + a)
    yn = orc.coset_lde_batch(np.ascontiguousarray(x.reshape(B, 1 << log_n)[:, rn]).ravel(), log_n, B, a,
                             orc.bb_to_monty(31)).reshape(B, 2 << log_n)[:, rN]
    d["lde_2e20x256_2x_rowmajor_shift31_natural"] = sha(yn.T)
    h = hashlib.sha256()
    for field, ln, seed in uneven_shapes():
        x = orc.lcg_fill(seed, 1 << ln, field)
        h.update(np.ascontiguousarray(orc.ntt_batch(field, False, ln, 1, x), dtype="<u4").tobytes())
    d["uneven_512_columns_forward"] = h.hexdigest()
    for k, (kind, log_n, batch, a, seed) in enumerate(serving_slice()):
        x = serving_slice_input(kind, log_n, batch, seed)
        d[f"serving_slice_{k}"] = sha(serving_slice_output(kind, log_n, batch, a, x))
    return d


if __name__ == "__main__":
    np.savez(os.path.join(HERE, "fixtures.npz"), **small_fixtures())
    with open(os.path.join(HERE, "digests.json"), "w") as fh:
        json.dump(digests(), fh, indent=1, sort_keys=True)
        fh.write("\n")
    print("wrote", os.listdir(HERE))
