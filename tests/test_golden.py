"""The oracle against the committed golden fixtures (tests/golden/, written by
tests/golden/make_golden.py): a regression pin of the CPU restatement, so a
change to it cannot silently move every GPU parity check with it."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M31, BB = orc.FIELD_M31, orc.FIELD_BB


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u4").tobytes()).hexdigest()


@pytest.fixture(scope="module")
def fx():
    with np.load(os.path.join(GOLDEN, "fixtures.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def digests():
    with open(os.path.join(GOLDEN, "digests.json")) as fh:
        return json.load(fh)


@pytest.mark.parametrize("name,field", [("bb", BB), ("m31", M31)])
@pytest.mark.parametrize("log_n", [4, 9])
def test_transforms_match_fixtures(fx, name, field, log_n):
    x = fx[f"{name}_fwd_{log_n}_in"]
    assert np.array_equal(x, orc.lcg_fill(777 + log_n, 1 << log_n, field))
    fwd = orc.m31_forward(x) if field == M31 else orc.bb_forward(x)
    inv = orc.m31_inverse(x) if field == M31 else orc.bb_inverse(x)
    assert np.array_equal(fwd, fx[f"{name}_fwd_{log_n}_out"])
    assert np.array_equal(inv, fx[f"{name}_inv_{log_n}_out"])


@pytest.mark.parametrize("log_n,batch,a", [(8, 4, 1), (8, 4, 2)])
def test_lde_matches_fixtures(fx, log_n, batch, a):
    x = fx[f"lde_{log_n}_{batch}_{a}_in"]
    assert np.array_equal(orc.coset_lde_batch(x, log_n, batch, a), fx[f"lde_{log_n}_{batch}_{a}_ref_out"])
    assert np.array_equal(orc.coset_lde_batch(x, log_n, batch, a, orc.bb_to_monty(31)),
                          fx[f"lde_{log_n}_{batch}_{a}_g31_out"])


def test_digests_of_single_column_transforms(digests):
    x = orc.lcg_fill(42, 8 << 12, BB)
    assert sha(orc.coset_lde_batch(x, 12, 8, 1)) == digests["lde_12_8_1_ref"]
    assert sha(orc.coset_lde_batch(x, 12, 8, 1, orc.bb_to_monty(31))) == digests["lde_12_8_1_g31"]
    n = 1 << 16
    assert sha(orc.bb_forward(orc.lcg_fill(12345 + n, n, BB))) == digests["bb_fwd_2e16"]
    n = 1 << 20
    assert sha(orc.m31_forward(orc.lcg_fill(12345 + n, n, M31))) == digests["m31_fwd_2e20"]


def test_digests_of_small_serving_slice_jobs(digests):
    # the slice jobs of at most 2^21 output elements, recomputed by the oracle
    from golden import make_golden as mg

    checked = 0
    for k, (kind, log_n, batch, a, seed) in enumerate(mg.serving_slice()):
        if (batch << (log_n + a)) > (1 << 21):
            continue
        x = mg.serving_slice_input(kind, log_n, batch, seed)
        assert sha(mg.serving_slice_output(kind, log_n, batch, a, x)) == digests[f"serving_slice_{k}"], k
        checked += 1
    assert checked >= 5
