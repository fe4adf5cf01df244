"""GPU: the allocation-free natural-order coset LDE
(sfn_coset_lde_natural_device_noalloc) against the allocating entry point
(sfn_coset_lde_natural_device) through the C ABI, bit-exact, on two calls in a
row (the second smaller), plus its refusals: tables not resident, output
beyond its 64 MiB static scratch, overlapping buffers."""
import ctypes

import numpy as np
import pytest

import closed_forms as cf
import oracle as orc
import sfntt
from sfntt import _lib as L

pytestmark = pytest.mark.gpu


def test_noalloc_natural_lde_matches_allocating_entry_point():
    ctx = sfntt.HipContext(0)
    lib = ctx._lib
    shift = orc.bb_to_monty(31)
    # row-major with tuned outer passes: the intermediate array is the static
    # scratch (2^16 x 64 at 2x = 32 MiB of output); then a smaller call; then SoA
    for log_n, batch, a, layout in ((16, 64, 1, L.LAYOUT_ROW_MAJOR), (12, 8, 2, L.LAYOUT_ROW_MAJOR),
                                    (13, 16, 1, L.LAYOUT_SOA)):
        rm = layout == L.LAYOUT_ROW_MAJOR
        x = orc.lcg_fill(808 + log_n, batch << log_n, orc.FIELD_BB)
        nout = batch << (log_n + a)
        din, dref, dna = sfntt.DeviceBuffer(ctx, x.size), sfntt.DeviceBuffer(ctx, nout), sfntt.DeviceBuffer(ctx, nout)
        din.upload(x)
        print(f"[noalloc] 2^{log_n} x {batch} a={a} layout={layout}", flush=True)
        ctx.check(lib.sfn_coset_lde_natural_device(ctx.ptr, log_n, batch, a, shift, layout, din.ptr, dref.ptr, None))
        ctx.check(lib.sfn_coset_lde_natural_device_noalloc(ctx.ptr, log_n, batch, a, shift, layout, din.ptr,
                                                           dna.ptr, None))
        ctx.check(lib.sfn_synchronize(ctx.ptr, None))
        ref, got = dref.download(), dna.download()
        assert np.array_equal(got, ref), f"allocation-free LDE 2^{log_n} x {batch} differs"
        exp = orc.coset_lde_batch(cf.rows_bitrev(x, log_n, batch, rm), log_n, batch, a, shift, layout=layout)
        assert np.array_equal(got, cf.rows_bitrev(exp, log_n + a, batch, rm)), "differs from the oracle"
        for b in (din, dref, dna):
            b.free()
    d = sfntt.DeviceBuffer(ctx, 48 << 20)
    base = d.ptr.value
    # a shape never run on this context: its tables are not resident
    assert lib.sfn_coset_lde_natural_device_noalloc(ctx.ptr, 11, 8, 1, shift, L.LAYOUT_ROW_MAJOR, d.ptr,
                                                    ctypes.c_void_p(base + (8 << 13)), None) == L.ERR_NOT_READY
    # 2^20 x 16 at 2x is 128 MiB of output: beyond the static scratch
    assert lib.sfn_coset_lde_natural_device_noalloc(ctx.ptr, 20, 16, 1, shift, L.LAYOUT_ROW_MAJOR, d.ptr,
                                                    ctypes.c_void_p(base + (64 << 20)), None) == L.ERR_INVALID_SIZE
    assert lib.sfn_coset_lde_natural_device_noalloc(ctx.ptr, 10, 4, 1, shift, L.LAYOUT_ROW_MAJOR, d.ptr,
                                                    ctypes.c_void_p(base + 4), None) == L.ERR_INVALID_ARGUMENT
    d.free()
    ctx.close()
