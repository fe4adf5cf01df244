"""GPU: every kernel path gives the same bits.

The library picks among several kernel families for one transform (tuned
radix-32 tiles, generic LDS-staged rounds, the XCD-grouped tile order, the
autotuner's schedules).  Each switch is a performance choice only: these tests
run the same inputs through a context with the switch off and compare the
outputs with each other and with the CPU oracle.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle as orc
import sfntt
from sfntt import _lib as L

pytestmark = pytest.mark.gpu

M31, BB, KB = sfntt.FIELD_M31, sfntt.FIELD_BABYBEAR, sfntt.FIELD_KOALABEAR


def context_with(env: dict):
    """A context created while `env` is set (the switches are read at creation)."""
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return sfntt.HipContext(0)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.fixture(scope="module")
def ctx():
    return sfntt.HipContext(0)


# (2^20 cases with batches of 8 and 16: row-major 2^10-point passes over
# column-fastest tiles, anchored on the oracle below)
CASES_NTT = [(M31, 12, 1), (M31, 17, 3), (M31, 20, 1), (M31, 20, 16), (BB, 11, 5), (BB, 16, 4), (BB, 20, 2),
             (BB, 20, 8), (KB, 13, 3), (KB, 20, 8)]
CASES_LDE = [(10, 24, 1), (14, 16, 2), (16, 8, 1), (20, 8, 1)]


def _run_all(c):
    out = []
    for field, log_n, batch in CASES_NTT:
        x = orc.lcg_fill(1000 + log_n, batch << log_n, field)
        for layout in (sfntt.LAYOUT_SOA, sfntt.LAYOUT_ROW_MAJOR):
            for d in (sfntt.FORWARD, sfntt.INVERSE):
                print(f"[paths] ntt field={field} log_n={log_n} batch={batch} layout={layout} dir={d}", flush=True)
                y, _ = c.ntt(field, sfntt.VARIANT_AUTO, d, log_n, batch, x, layout)
                out.append(y)
    for log_n, batch, a in CASES_LDE:
        x = orc.lcg_fill(42 + log_n, batch << log_n, BB)
        for layout, shift in ((sfntt.LAYOUT_SOA, 0), (sfntt.LAYOUT_ROW_MAJOR, orc.bb_to_monty(31))):
            print(f"[paths] lde log_n={log_n} batch={batch} a={a} layout={layout}", flush=True)
            y, _ = c.coset_lde(x, log_n, batch, a, shift, layout)
            out.append(y)
    return out


@pytest.fixture(scope="module")
def reference_outputs(ctx):
    outs = _run_all(ctx)
    # anchor the default path on the oracle for a few of the cases
    x = orc.lcg_fill(1000 + 16, 4 << 16, BB)
    exp = orc.ntt_batch(BB, False, 16, 4, x)
    assert np.array_equal(outs[4 * CASES_NTT.index((BB, 16, 4))], exp), "default path differs from the oracle"
    for i, (field, log_n, batch) in enumerate(CASES_NTT):
        if log_n == 20 and batch >= 8:  # row-major outputs
            x = orc.lcg_fill(1000 + log_n, batch << log_n, field)
            for d in (0, 1):
                exp = orc.ntt_batch(field, d == 1, log_n, batch, x, layout=orc.LAYOUT_ROW_MAJOR)
                assert np.array_equal(outs[4 * i + 2 + d], exp), f"row-major field={field} batch={batch} dir={d}"
    for i, (field, log_n, batch) in enumerate(CASES_NTT):
        if field == KB:  # KoalaBear: forward and inverse, SoA
            x = orc.lcg_fill(1000 + log_n, batch << log_n, KB)
            assert np.array_equal(outs[4 * i], orc.ntt_batch(KB, False, log_n, batch, x)), "KoalaBear forward"
            assert np.array_equal(outs[4 * i + 1], orc.ntt_batch(KB, True, log_n, batch, x)), "KoalaBear inverse"
    x = orc.lcg_fill(42 + 14, 16 << 14, BB)
    assert np.array_equal(outs[4 * len(CASES_NTT) + 2], orc.coset_lde_batch(x, 14, 16, 2)), "default LDE differs"
    x = orc.lcg_fill(42 + 20, 8 << 20, BB)
    exp = orc.coset_lde_batch(x, 20, 8, 1, orc.bb_to_monty(31), orc.LAYOUT_ROW_MAJOR)
    assert np.array_equal(outs[4 * len(CASES_NTT) + 2 * 3 + 1], exp), "row-major 2^20 x 8 LDE differs"
    return outs


@pytest.mark.parametrize("env", [{"SFNTT_GENERIC_ONLY": "1"}, {"SFNTT_XCD_REMAP": "0"}],
                         ids=["generic_kernels", "xcd_order_off"])
def test_switch_gives_same_bits(reference_outputs, env):
    c = context_with(env)
    got = _run_all(c)
    assert len(got) == len(reference_outputs)
    for i, (g, r) in enumerate(zip(got, reference_outputs)):
        assert np.array_equal(g, r), f"case {i} differs with {env}"
    c.close()


def test_autotuned_schedules_give_same_bits(ctx, reference_outputs):
    c = sfntt.HipContext(0)
    best = ctypes.c_uint32()
    for field, log_n, batch in CASES_NTT:
        for layout in (sfntt.LAYOUT_SOA, sfntt.LAYOUT_ROW_MAJOR):
            c.check(c._lib.sfn_autotune(c.ptr, field, log_n, batch, layout, ctypes.byref(best)))
    for log_n, batch, _ in CASES_LDE:  # the LDE (2x)
        for layout in (sfntt.LAYOUT_SOA, sfntt.LAYOUT_ROW_MAJOR):
            c.check(c._lib.sfn_autotune(c.ptr, sfntt._lib.TUNE_LDE, log_n, batch, layout, ctypes.byref(best)))
    got = _run_all(c)
    for i, (g, r) in enumerate(zip(got, reference_outputs)):
        assert np.array_equal(g, r), f"case {i} differs under autotuned schedules"
    c.close()


def autotune_candidates(log_n, lde):
    """The schedules the autotuner times for one shape (batch.cpp, sfn_autotune):
    (largest tile, register radix, lanes per workgroup, middle-pass lanes)."""
    for mlt in range(6, 12):
        if (log_n + mlt - 1) // mlt > 3 or (lde and log_n > 2 * mlt):
            continue
        for logr in (3, 4, 5):
            for lw in (2, 3, 4, 5):
                for mlw in ((2, 3, 4, 5) if lde else (lw,)):
                    if mlt + max(lw, mlw) <= 15:
                        yield mlt, logr, lw, mlw


def _force_schedule(c, key_field, log_n, batch, layout, sched):
    mlt, logr, lw, mlw = sched
    js = (f'[{{"field":{key_field},"log_n":{log_n},"batch":{batch},"layout":{layout},'
          f'"max_log_t":{mlt},"logr":{logr},"log_w":{lw},"mid_log_w":{mlw}}}]')
    c.check(c._lib.sfn_tuning_load(c.ptr, js.encode()))


NOT_LAUNCHABLE = (sfntt.api.Unsupported, sfntt.api.InvalidArgument, sfntt.api.InvalidSize)


@pytest.mark.parametrize("field,log_n,batch,layout", [
    (BB, 16, 4, sfntt.LAYOUT_SOA), (BB, 16, 4, sfntt.LAYOUT_ROW_MAJOR), (M31, 17, 3, sfntt.LAYOUT_SOA),
    (KB, 13, 3, sfntt.LAYOUT_ROW_MAJOR)])
def test_every_autotuner_schedule_matches_oracle(field, log_n, batch, layout):
    """Each candidate the autotuner may pick gives the oracle's bits, forward and inverse
    (the autotuner itself only times them)."""
    c = sfntt.HipContext(0)
    x = orc.lcg_fill(1000 + log_n, batch << log_n, field)
    exp = [orc.ntt_batch(field, d == 1, log_n, batch, x, layout=layout) for d in (0, 1)]
    ran = 0
    for sched in autotune_candidates(log_n, False):
        print(f"[sched] field={field} log_n={log_n} batch={batch} layout={layout} schedule={sched}", flush=True)
        _force_schedule(c, field, log_n, batch, layout, sched)
        for d in (sfntt.FORWARD, sfntt.INVERSE):
            try:
                y, _ = c.ntt(field, sfntt.VARIANT_AUTO, d, log_n, batch, x, layout)
            except NOT_LAUNCHABLE:
                continue
            assert np.array_equal(y, exp[d]), f"schedule {sched} dir {d}"
            ran += 1
    assert ran >= 20
    c.close()


@pytest.mark.parametrize("log_n,batch,layout", [(12, 6, sfntt.LAYOUT_SOA), (13, 8, sfntt.LAYOUT_ROW_MAJOR)])
def test_every_autotuner_lde_schedule_matches_oracle(log_n, batch, layout):
    """Each LDE candidate (outer and middle-pass lanes tuned apart) gives the oracle's bits."""
    c = sfntt.HipContext(0)
    x = orc.lcg_fill(42 + log_n, batch << log_n, BB)
    shift = 0 if layout == sfntt.LAYOUT_SOA else orc.bb_to_monty(31)
    exp = orc.coset_lde_batch(x, log_n, batch, 1, shift, layout)
    ran = 0
    for sched in autotune_candidates(log_n, True):
        print(f"[sched] lde log_n={log_n} batch={batch} layout={layout} schedule={sched}", flush=True)
        _force_schedule(c, L.TUNE_LDE, log_n, batch, layout, sched)
        try:
            y, _ = c.coset_lde(x, log_n, batch, 1, shift, layout)
        except NOT_LAUNCHABLE:
            continue
        assert np.array_equal(y, exp), f"schedule {sched}"
        ran += 1
    assert ran >= 20
    c.close()
