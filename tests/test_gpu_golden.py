"""GPU outputs against the committed golden fixtures and the SHA-256 digests of
the full benchmark workloads (tests/golden/, written from the oracle by
tests/golden/make_golden.py).  Full sizes: BabyBear forward 2^16 and M31
forward 2^20 (all variants), the coset LDE 2^20 x 256 2x (row-major shift 31
and SoA reference shift), 512 uneven columns through a graph plan."""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as orc
import sfntt
from sfntt import _lib as L

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M31, BB = sfntt.FIELD_M31, sfntt.FIELD_BABYBEAR


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u4").tobytes()).hexdigest()


@pytest.fixture(scope="module")
def ctx():
    return sfntt.HipContext(0)


@pytest.fixture(scope="module")
def fx():
    with np.load(os.path.join(GOLDEN, "fixtures.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def digests():
    with open(os.path.join(GOLDEN, "digests.json")) as fh:
        return json.load(fh)


def test_small_fixtures(ctx, fx):
    for name, field in (("bb", BB), ("m31", M31)):
        for log_n in (4, 9):
            x = fx[f"{name}_fwd_{log_n}_in"]
            for v in (sfntt.CT_DIT_R2, sfntt.CT_GS_R2, sfntt.STOCKHAM_R2, sfntt.CT_GS_R4, sfntt.VARIANT_AUTO):
                y, _ = ctx.ntt(field, v, sfntt.FORWARD, log_n, 1, x)
                assert np.array_equal(y, fx[f"{name}_fwd_{log_n}_out"]), f"{name} fwd 2^{log_n} v{v}"
            y, _ = ctx.ntt(field, sfntt.VARIANT_AUTO, sfntt.INVERSE, log_n, 1, x)
            assert np.array_equal(y, fx[f"{name}_inv_{log_n}_out"]), f"{name} inv 2^{log_n}"
    for log_n, batch, a in ((8, 4, 1), (8, 4, 2)):
        x = fx[f"lde_{log_n}_{batch}_{a}_in"]
        y, _ = ctx.coset_lde(x, log_n, batch, a)
        assert np.array_equal(y, fx[f"lde_{log_n}_{batch}_{a}_ref_out"])
        y, _ = ctx.coset_lde(x, log_n, batch, a, orc.bb_to_monty(31))
        assert np.array_equal(y, fx[f"lde_{log_n}_{batch}_{a}_g31_out"])


def test_single_column_digests(ctx, digests):
    n = 1 << 16
    x = orc.lcg_fill(12345 + n, n, BB)
    for v in (sfntt.CT_DIT_R2, sfntt.CT_GS_R2, sfntt.STOCKHAM_R2, sfntt.CT_GS_R4, sfntt.VARIANT_AUTO):
        y, _ = ctx.ntt(BB, v, sfntt.FORWARD, 16, 1, x)
        assert sha(y) == digests["bb_fwd_2e16"], f"BabyBear 2^16 v{v}"
    n = 1 << 20
    x = orc.lcg_fill(12345 + n, n, M31)
    for v in (sfntt.CT_DIT_R2, sfntt.CT_GS_R2, sfntt.STOCKHAM_R2, sfntt.CT_GS_R4, sfntt.VARIANT_AUTO):
        print(f"[golden] m31 2^20 v{v}", flush=True)
        y, _ = ctx.ntt(M31, v, sfntt.FORWARD, 20, 1, x)
        assert sha(y) == digests["m31_fwd_2e20"], f"M31 2^20 v{v}"
    x = orc.lcg_fill(42, 8 << 12, BB)
    y, _ = ctx.coset_lde(x, 12, 8, 1)
    assert sha(y) == digests["lde_12_8_1_ref"]
    y, _ = ctx.coset_lde(x, 12, 8, 1, orc.bb_to_monty(31))
    assert sha(y) == digests["lde_12_8_1_g31"]


def test_lde_2e20x256_digests(ctx, digests):
    log_n, B, a = 20, 256, 1
    n, N = 1 << log_n, 2 << log_n
    x = orc.lcg_fill(42 + log_n, B << log_n, BB)  # generated column by column
    d_in, d_out = sfntt.DeviceBuffer(ctx, n * B), sfntt.DeviceBuffer(ctx, N * B)
    try:
        print("[golden] lde 2^20 x 256 row-major shift 31", flush=True)
        d_in.upload(x.reshape(B, n).T.copy().ravel())
        ctx.check(ctx._lib.sfn_coset_lde_batch_device(ctx.ptr, log_n, B, a, orc.bb_to_monty(31),
                                                      L.LAYOUT_ROW_MAJOR, d_in.ptr, d_out.ptr, None))
        ctx.check(ctx._lib.sfn_synchronize(ctx.ptr, None))
        assert sha(d_out.download()) == digests["lde_2e20x256_2x_rowmajor_shift31"]
        print("[golden] lde 2^20 x 256 SoA reference shift", flush=True)
        d_in.upload(x)
        ctx.check(ctx._lib.sfn_coset_lde_batch_device(ctx.ptr, log_n, B, a, 0, L.LAYOUT_SOA, d_in.ptr,
                                                      d_out.ptr, None))
        ctx.check(ctx._lib.sfn_synchronize(ctx.ptr, None))
        assert sha(d_out.download()) == digests["lde_2e20x256_2x_soa_refshift"]
        print("[golden] natural-order lde 2^20 x 256 row-major shift 31", flush=True)
        d_in.upload(x.reshape(B, n).T.copy().ravel())  # the same columns as natural-order rows
        ctx.check(ctx._lib.sfn_coset_lde_natural_device(ctx.ptr, log_n, B, a, orc.bb_to_monty(31),
                                                        L.LAYOUT_ROW_MAJOR, d_in.ptr, d_out.ptr, None))
        ctx.check(ctx._lib.sfn_synchronize(ctx.ptr, None))
        assert sha(d_out.download()) == digests["lde_2e20x256_2x_rowmajor_shift31_natural"]
    finally:
        d_in.free()
        d_out.free()


def test_uneven_512_columns_plan_digest(ctx, digests):
    it = orc.splitmix64_stream(0x5EED0004)
    shapes = []
    for _ in range(512):
        a, b, c = next(it), next(it), next(it)
        shapes.append((int(a & 1), 10 + int(b % 13), int(c)))
    bufs, descs = [], []
    for field, log_n, seed in shapes:
        buf = sfntt.DeviceBuffer(ctx, 1 << log_n)
        buf.upload(orc.lcg_fill(seed, 1 << log_n, field))
        bufs.append(buf)
        descs.append(L.JobDesc(field, sfntt.FORWARD, log_n, 1, buf.ptr, buf.ptr))
    arr = (L.JobDesc * len(descs))(*descs)
    plan = ctypes.c_void_p()
    ctx.check(ctx._lib.sfn_plan_create(ctx.ptr, arr, len(descs), 1, ctypes.byref(plan)))
    try:
        print("[golden] uneven 512 columns plan (graph)", flush=True)
        ctx.check(ctx._lib.sfn_plan_launch(plan, None))
        ctx.check(ctx._lib.sfn_synchronize(ctx.ptr, None))
        h = hashlib.sha256()
        for b in bufs:
            h.update(np.ascontiguousarray(b.download(), dtype="<u4").tobytes())
        assert h.hexdigest() == digests["uneven_512_columns_forward"]
    finally:
        ctx._lib.sfn_plan_destroy(plan)
        for b in bufs:
            b.free()


def test_serving_slice_at_stated_sizes(ctx, digests):
    # 20 jobs of the serving stream at its stated sizes (a 2^20 x 256 LDE at 4x
    # and at 2x, an M31 2^22 NTT, 17 drawn jobs) submitted to one server at
    # once, even jobs from host memory and odd jobs from device memory; every
    # output against the oracle's digest (src/lde/mod.rs:194-271 compares the
    # GPU LDE with the CPU one the same way, per job)
    from golden import make_golden as mg

    lib = ctx._lib
    cfg = L.ServerConfig(0, 4, 0, 0, 0)
    srv = ctypes.c_void_p()
    ctx.check(lib.sfn_server_create(ctx.ptr, ctypes.byref(cfg), ctypes.byref(srv)))
    keep = []
    try:
        for k, (kind, log_n, batch, a, seed) in enumerate(mg.serving_slice()):
            x = mg.serving_slice_input(kind, log_n, batch, seed)
            nout = batch << (log_n + a)
            dev = k % 2
            print(f"[golden] serving slice job {k}: {kind} 2^{log_n} x {batch} a={a} device_ptrs={dev}", flush=True)
            if dev:
                din, dout = sfntt.DeviceBuffer(ctx, x.size), sfntt.DeviceBuffer(ctx, nout)
                din.upload(x)
                inp, outp, out = din.ptr, dout.ptr, dout
            else:
                out = np.zeros(nout, np.uint32)
                inp, outp, din = x.ctypes.data, out.ctypes.data, x
            r = L.Request(L.JOB_LDE if kind == "lde" else L.JOB_NTT, BB if kind == "lde" else M31, L.FORWARD,
                          log_n, batch, a, 0, dev, inp, outp)
            t = ctypes.c_uint64()
            ctx.check(lib.sfn_server_submit(srv, ctypes.byref(r), ctypes.byref(t)))
            keep.append((k, t.value, dev, din, out))
        for k, t, dev, din, out in keep:
            tm = L.JobTiming()
            ctx.check(lib.sfn_server_wait(srv, t, ctypes.byref(tm)))
            got = out.download() if dev else out
            assert sha(got) == digests[f"serving_slice_{k}"], f"serving slice job {k}"
            if dev:
                din.free()
                out.free()
    finally:
        lib.sfn_server_destroy(srv)
