"""GPU tests of the multi-job plan (uneven sizes, both fields, one launch per
pass, hipGraph replay), the autotuner, and the serving loop (every result the
same alone and in the mix).  Expected values come from the CPU oracle, which
restates each job's single-column reference transform."""
import ctypes
import threading
import time

import numpy as np
import pytest

import oracle as orc
import sfntt
from sfntt import _lib as L
from sfntt import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return sfntt.HipContext(0)


def uneven_jobs(seed, count, lo, hi):
    """Uneven-columns draw: field ~ Bernoulli(1/2), log_n ~ U{lo..hi}, per-column LCG seed."""
    st = orc.splitmix64_stream(seed)
    jobs = []
    for _ in range(count):
        f = next(st) & 1
        log_n = lo + next(st) % (hi - lo + 1)
        s = next(st)
        jobs.append((int(f), int(log_n), int(s)))
    return jobs


def run_plan(ctx, specs, use_graph, dirs=None, reps=1):
    bufs, descs, inputs = [], [], []
    for k, (f, log_n, s) in enumerate(specs):
        x = synth.field_data(f, s, 1 << log_n)
        b = sfntt.DeviceBuffer(ctx, x.size)
        b.upload(x)
        bufs.append(b)
        inputs.append(x)
        d = L.JobDesc(f, dirs[k] if dirs else L.FORWARD, log_n, 1, b.ptr, b.ptr)
        descs.append(d)
    arr = (L.JobDesc * len(descs))(*descs)
    plan = ctypes.c_void_p()
    ctx.check(ctx._lib.sfn_plan_create(ctx.ptr, arr, len(descs), int(use_graph), ctypes.byref(plan)))
    nl, nt = ctypes.c_uint32(), ctypes.c_uint32()
    ctx.check(ctx._lib.sfn_plan_info(plan, ctypes.byref(nl), ctypes.byref(nt)))
    for _ in range(reps):
        ctx.check(ctx._lib.sfn_plan_launch(plan, None))
    ctx.check(ctx._lib.sfn_synchronize(ctx.ptr, None))
    outs = [b.download() for b in bufs]
    ctx._lib.sfn_plan_destroy(plan)
    return inputs, outs, nl.value


@pytest.mark.parametrize("use_graph", [0, 1], ids=["direct", "graph"])
def test_uneven_mixed_plan_matches_single_columns(ctx, use_graph):
    specs = uneven_jobs(0x5EED0004, 48, 10, 18)
    dirs = [k % 2 for k in range(len(specs))]
    inputs, outs, launches = run_plan(ctx, specs, use_graph, dirs)
    assert launches <= 3  # one launch per pass, not per job
    for (f, log_n, s), d, x, y in zip(specs, dirs, inputs, outs):
        fn = {(0, 0): orc.m31_forward, (0, 1): orc.m31_inverse,
              (1, 0): orc.bb_forward, (1, 1): orc.bb_inverse}[(f, d)]
        assert np.array_equal(y, fn(x)), (f, log_n, d)


@pytest.mark.parametrize("in_place", [True, False], ids=["in_place", "out_of_place"])
def test_plan_of_contiguous_small_columns(ctx, in_place):
    # Many small single-column jobs laid out back to back in one buffer (the
    # batched small-NTT case), with runs of changing shape and direction: the
    # plan merges each run into one batched job; every column must still equal
    # its own single-column transform.
    runs = [(orc.FIELD_BB, 8, 300, 0), (orc.FIELD_M31, 8, 7, 0), (orc.FIELD_BB, 8, 50, 1), (orc.FIELD_BB, 10, 33, 0),
            (orc.FIELD_M31, 10, 40, 1), (orc.FIELD_BB, 6, 5, 0)]
    total = sum(c << ln for _, ln, c, _ in runs)
    src = sfntt.DeviceBuffer(ctx, total)
    dst = src if in_place else sfntt.DeviceBuffer(ctx, total)
    xs, descs, off = [], [], 0
    for f, ln, cols, d in runs:
        for c in range(cols):
            x = orc.lcg_fill(off + 17, 1 << ln, f)
            xs.append((f, ln, d, off, x))
            descs.append(L.JobDesc(f, d, ln, 1, ctypes.c_void_p(src.ptr.value + 4 * off),
                                   ctypes.c_void_p(dst.ptr.value + 4 * off)))
            off += 1 << ln
    host = np.concatenate([x for *_, x in xs])
    src.upload(host)
    arr = (L.JobDesc * len(descs))(*descs)
    plan = ctypes.c_void_p()
    ctx.check(ctx._lib.sfn_plan_create(ctx.ptr, arr, len(descs), 1, ctypes.byref(plan)))
    ctx.check(ctx._lib.sfn_plan_launch(plan, None))
    ctx.check(ctx._lib.sfn_synchronize(ctx.ptr, None))
    ctx._lib.sfn_plan_destroy(plan)
    y = dst.download()
    for f, ln, d, o, x in xs:
        exp = orc.ntt_batch(f, d == 1, ln, 1, x)
        assert np.array_equal(y[o:o + (1 << ln)], exp), f"column at {o} field={f} log_n={ln} dir={d}"


@pytest.mark.parametrize("mix", ["koalabear_only", "all_fields"])
def test_plan_with_koalabear_jobs(ctx, mix):
    # KoalaBear jobs alone run on the tuned many-job kernel with KoalaBear's
    # root table; mixed with BabyBear (two root tables) the plan takes the
    # generic many-job kernel.  Every column equals its single-column transform.
    st = orc.splitmix64_stream(0x6B0A1A)
    specs, dirs = [], []
    for k in range(40):
        f = 2 if mix == "koalabear_only" else int(next(st) % 3)
        specs.append((f, int(8 + next(st) % 13), int(next(st))))
        dirs.append(k % 2)
    inputs, outs, launches = run_plan(ctx, specs, 1, dirs)
    assert launches <= 3
    for (f, log_n, s), d, x, y in zip(specs, dirs, inputs, outs):
        assert np.array_equal(y, orc.ntt_batch(f, d == 1, log_n, 1, x)), (f, log_n, d)


def test_graph_replays_each_match(ctx):
    # replay several times with host work between replays; every replay is a
    # fresh forward of the previous output, compared with the oracle each time
    specs = uneven_jobs(77, 12, 10, 16)
    bufs, descs, cur = [], [], []
    for f, log_n, s in specs:
        x = synth.field_data(f, s, 1 << log_n)
        b = sfntt.DeviceBuffer(ctx, x.size)
        b.upload(x)
        bufs.append(b)
        cur.append(x)
        descs.append(L.JobDesc(f, L.FORWARD, log_n, 1, b.ptr, b.ptr))
    arr = (L.JobDesc * len(descs))(*descs)
    plan = ctypes.c_void_p()
    ctx.check(ctx._lib.sfn_plan_create(ctx.ptr, arr, len(descs), 1, ctypes.byref(plan)))
    for rep in range(3):
        ctx.check(ctx._lib.sfn_plan_launch(plan, None))
        ctx.check(ctx._lib.sfn_synchronize(ctx.ptr, None))
        for k, ((f, log_n, s), b) in enumerate(zip(specs, bufs)):
            exp = orc.m31_forward(cur[k]) if f == 0 else orc.bb_forward(cur[k])
            got = b.download()
            assert np.array_equal(got, exp), (rep, k)
            cur[k] = exp
    ctx._lib.sfn_plan_destroy(plan)


def test_uneven_512_columns_full_size_plan(ctx):
    # 512 columns, log_n ~ U{10..22}, both fields (the workload as stated)
    specs = uneven_jobs(0x5EED0004, 512, 10, 22)
    inputs, outs, launches = run_plan(ctx, specs, 1)
    rng = np.random.default_rng(4)
    check = rng.choice(len(specs), 40, replace=False)
    for k in check:
        f, log_n, s = specs[k]
        exp = orc.m31_forward(inputs[k]) if f == 0 else orc.bb_forward(inputs[k])
        assert np.array_equal(outs[k], exp), (f, log_n)


def test_autotune_and_tuning_roundtrip(ctx):
    best = ctypes.c_uint32()
    ctx.check(ctx._lib.sfn_autotune(ctx.ptr, L.FIELD_BABYBEAR, 16, 4, L.LAYOUT_SOA, ctypes.byref(best)))
    assert 6 <= best.value <= 11
    n = ctypes.c_size_t()
    ctx.check(ctx._lib.sfn_tuning_save(ctx.ptr, None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value)
    ctx.check(ctx._lib.sfn_tuning_save(ctx.ptr, buf, n.value, ctypes.byref(n)))
    assert b'"log_n":16' in buf.value
    ctx.check(ctx._lib.sfn_tuning_load(ctx.ptr, buf.value))
    # the tuned plan still gives the reference result
    x = orc.lcg_fill(5, 4 << 16, orc.FIELD_BB)
    y, _ = ctx.ntt(L.FIELD_BABYBEAR, L.VARIANT_AUTO, L.FORWARD, 16, 4, x)
    assert np.array_equal(y, orc.ntt_batch(1, False, 16, 4, x))
    with pytest.raises(sfntt.api.InvalidArgument):
        ctx.check(ctx._lib.sfn_tuning_load(ctx.ptr, b'[{"field":1}]'))


def test_call_timing_is_opt_in(ctx):
    # device-pointer calls record the per-call event pair only with sfn_call_timing on;
    # host-pointer calls always time themselves
    lib = ctx._lib
    log_n = 12
    x = orc.lcg_fill(77, 1 << log_n, orc.FIELD_BB)
    exp = orc.ntt_batch(orc.FIELD_BB, False, log_n, 1, x)
    din, dout = sfntt.DeviceBuffer(ctx, 1 << log_n), sfntt.DeviceBuffer(ctx, 1 << log_n)
    din.upload(x)
    ns = ctypes.c_uint64()

    def device_call():
        ctx.check(lib.sfn_ntt_device(ctx.ptr, L.FIELD_BABYBEAR, L.VARIANT_AUTO, L.FORWARD, log_n, 1,
                                     L.LAYOUT_SOA, din.ptr, dout.ptr, None))

    device_call()
    with pytest.raises(sfntt.api.InvalidArgument):
        ctx.check(lib.sfn_last_device_ns(ctx.ptr, ctypes.byref(ns)))
    assert np.array_equal(dout.download(), exp)
    ctx.check(lib.sfn_call_timing(ctx.ptr, 1))
    try:
        device_call()
        ctx.check(lib.sfn_last_device_ns(ctx.ptr, ctypes.byref(ns)))
        assert 0 < ns.value < 10**9
    finally:
        ctx.check(lib.sfn_call_timing(ctx.ptr, 0))
    assert np.array_equal(dout.download(), exp)
    y, dev_ns = ctx.ntt(L.FIELD_BABYBEAR, L.VARIANT_AUTO, L.FORWARD, log_n, 1, x)
    ctx.check(lib.sfn_last_device_ns(ctx.ptr, ctypes.byref(ns)))
    assert np.array_equal(y, exp) and ns.value == dev_ns > 0
    din.free()
    dout.free()


# ------------------------------------------------------------------ serving
def make_server(ctx, **kw):
    cfg = L.ServerConfig(kw.get("mem", 0), kw.get("inflight", 4), kw.get("queue", 0),
                         kw.get("group", 0), kw.get("cols", 0))
    srv = ctypes.c_void_p()
    ctx.check(ctx._lib.sfn_server_create(ctx.ptr, ctypes.byref(cfg), ctypes.byref(srv)))
    return srv


def mixed_requests(seed, count):
    """Mixed serving trace (scaled down): 60% BabyBear LDE, 40% M31 NTT."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        if rng.random() < 0.6:
            out.append(("lde", int(rng.integers(8, 13)), int(rng.integers(1, 9)), int(rng.integers(1, 3))))
        else:
            out.append(("ntt", int(rng.integers(10, 15)), 1, 0))
    return out


def expected(kind, log_n, batch, a, x):
    if kind == "lde":
        return orc.coset_lde_batch(x, log_n, batch, a)
    return orc.m31_forward(x)


@pytest.mark.parametrize("device_ptrs", [0, 1], ids=["host", "hbm"])
def test_serving_results_same_alone_and_in_mix(ctx, device_ptrs):
    reqs = mixed_requests(5, 60)
    srv = make_server(ctx, inflight=4)
    keep, tickets = [], []
    for i, (kind, log_n, batch, a) in enumerate(reqs):
        f = 1 if kind == "lde" else 0
        x = synth.field_data(f, 1000 + i, batch << log_n)
        nout = batch << (log_n + a)
        if device_ptrs:
            din, dout = sfntt.DeviceBuffer(ctx, x.size), sfntt.DeviceBuffer(ctx, nout)
            din.upload(x)
            inp, outp, out = din.ptr, dout.ptr, dout
        else:
            out = np.zeros(nout, np.uint32)
            inp, outp = x.ctypes.data, out.ctypes.data
            din = None
        r = L.Request(L.JOB_LDE if kind == "lde" else L.JOB_NTT, f, L.FORWARD, log_n, batch, a, 0,
                      device_ptrs, inp, outp)
        t = ctypes.c_uint64()
        ctx.check(ctx._lib.sfn_server_submit(srv, ctypes.byref(r), ctypes.byref(t)))
        keep.append((kind, log_n, batch, a, x, out, din))
        tickets.append(t.value)
    groups = []
    for t in tickets:
        tm = L.JobTiming()
        ctx.check(ctx._lib.sfn_server_wait(srv, t, ctypes.byref(tm)))
        assert tm.submit_ns <= tm.start_ns <= tm.end_ns
        groups.append(tm.group_size)
    ctx._lib.sfn_server_destroy(srv)
    for kind, log_n, batch, a, x, out, din in keep:
        got = out.download() if device_ptrs else out
        assert np.array_equal(got, expected(kind, log_n, batch, a, x)), (kind, log_n, batch, a)


def test_serving_koalabear_ntt_jobs(ctx):
    # KoalaBear NTT jobs of both directions through the serving loop (same-shape
    # jobs are grouped into one many-job launch per pass); a KoalaBear LDE is
    # refused (the coset LDE is BabyBear's) and 2^25 exceeds the two-adicity.
    srv = make_server(ctx, inflight=2)
    keep = []
    for i in range(24):
        log_n, d = 10 + i % 5, i % 2
        x = synth.field_data(2, 300 + i, 1 << log_n)
        out = np.zeros_like(x)
        r = L.Request(L.JOB_NTT, L.FIELD_KOALABEAR, d, log_n, 1, 0, 0, 0, x.ctypes.data, out.ctypes.data)
        t = ctypes.c_uint64()
        ctx.check(ctx._lib.sfn_server_submit(srv, ctypes.byref(r), ctypes.byref(t)))
        keep.append((t.value, log_n, d, x, out))
    for t, log_n, d, x, out in keep:
        ctx.check(ctx._lib.sfn_server_wait(srv, t, None))
        assert np.array_equal(out, orc.ntt_batch(2, d == 1, log_n, 1, x)), (log_n, d)
    tb = ctypes.c_uint64()
    x = np.zeros(1 << 10, np.uint32)
    lde = L.Request(L.JOB_LDE, L.FIELD_KOALABEAR, 0, 10, 1, 1, 0, 0, x.ctypes.data, x.ctypes.data)
    assert ctx._lib.sfn_server_submit(srv, ctypes.byref(lde), ctypes.byref(tb)) == L.ERR_UNSUPPORTED
    big = L.Request(L.JOB_NTT, L.FIELD_KOALABEAR, 0, 25, 1, 0, 0, 0, x.ctypes.data, x.ctypes.data)
    assert ctx._lib.sfn_server_submit(srv, ctypes.byref(big), ctypes.byref(tb)) == L.ERR_INVALID_SIZE
    ctx._lib.sfn_server_destroy(srv)


def test_server_and_owner_thread_share_tables():
    # The server's workers and the owning thread use one context at once: the
    # workers run small NTTs (log_n 3..5 take the generic kernels and their
    # position tables) while the owner runs generic variants at growing sizes,
    # so the position tables grow under the workers, and re-tunes a shape.
    c = sfntt.HipContext(0)
    srv = make_server(c, inflight=4)
    keep = []
    for i in range(64):
        f = (1, 2, 0)[i % 3]
        log_n, d = 3 + i % 9, (i // 3) % 2
        x = synth.field_data(f, 700 + i, 1 << log_n)
        out = np.zeros_like(x)
        r = L.Request(L.JOB_NTT, f, d, log_n, 1, 0, 0, 0, x.ctypes.data, out.ctypes.data)
        t = ctypes.c_uint64()
        c.check(c._lib.sfn_server_submit(srv, ctypes.byref(r), ctypes.byref(t)))
        keep.append((t.value, f, log_n, d, x, out))
    best = ctypes.c_uint32()
    for log_n in (12, 14, 16, 18, 20):
        for f in (L.FIELD_BABYBEAR, L.FIELD_KOALABEAR):
            x = synth.field_data(f, 900 + log_n, 1 << log_n)
            for v in (L.CT_GS_R2, L.CT_GS_R4):
                y, _ = c.ntt(f, v, L.FORWARD, log_n, 1, x)
                assert np.array_equal(y, orc.ntt_batch(f, False, log_n, 1, x)), (f, v, log_n)
        c.check(c._lib.sfn_autotune(c.ptr, L.FIELD_BABYBEAR, 8, 1, L.LAYOUT_SOA, ctypes.byref(best)))
    for t, f, log_n, d, x, out in keep:
        c.check(c._lib.sfn_server_wait(srv, t, None))
        assert np.array_equal(out, orc.ntt_batch(f, d == 1, log_n, 1, x)), (f, log_n, d)
    c._lib.sfn_server_destroy(srv)
    c.close()


def test_serving_admission_and_errors(ctx):
    srv = make_server(ctx, inflight=1, mem=1 << 20, queue=2)
    big = np.zeros(1 << 20, np.uint32)
    out = np.zeros(1 << 20, np.uint32)
    t = ctypes.c_uint64()
    r = L.Request(L.JOB_NTT, 0, 0, 20, 1, 0, 0, 0, big.ctypes.data, out.ctypes.data)
    # 8 MiB footprint never fits a 1 MiB budget: refused at submit
    assert ctx._lib.sfn_server_submit(srv, ctypes.byref(r), ctypes.byref(t)) == L.ERR_BUFFER_ALLOC_FAILED
    bad = L.Request(L.JOB_LDE, 0, 0, 10, 1, 1, 0, 0, big.ctypes.data, out.ctypes.data)
    assert ctx._lib.sfn_server_submit(srv, ctypes.byref(bad), ctypes.byref(t)) == L.ERR_UNSUPPORTED
    bad2 = L.Request(L.JOB_LDE, 1, 0, 26, 1, 1, 0, 0, big.ctypes.data, out.ctypes.data)
    assert ctx._lib.sfn_server_submit(srv, ctypes.byref(bad2), ctypes.byref(t)) == L.ERR_INVALID_SIZE
    assert ctx._lib.sfn_server_poll(srv, 12345, None) == L.ERR_INVALID_ARGUMENT
    ctx._lib.sfn_server_destroy(srv)


def test_table_cache_stays_within_its_budget():
    # Every LDE shift has its own post / point-factor tables: 30 distinct shifts
    # under a 1 MiB budget keep the cache bounded, results stay exact, and a
    # plan's tables survive evictions until the plan is destroyed
    # (the reference's TwiddleCache keeps one log_n, src/ntt/twiddle_cache.rs:29-82).
    c = sfntt.HipContext(0)
    x = orc.lcg_fill(3, 8 << 12, orc.FIELD_BB)
    c.coset_lde(x, 12, 8, 1, orc.bb_to_monty(3))
    base = c.twiddles_bytes()
    c.twiddles_limit(1 << 20)
    for k in range(30):
        shift = orc.bb_to_monty(5 + 2 * k)
        y, _ = c.coset_lde(x, 12, 8, 1, shift)
        assert np.array_equal(y, orc.coset_lde_batch(x, 12, 8, 1, shift)), k
        assert c.twiddles_bytes() <= base + (1 << 20), (k, c.twiddles_bytes())
    specs = [(0, 13, 11), (1, 14, 12), (0, 18, 13), (1, 16, 14)]
    bufs, descs, xs = [], [], []
    for f, log_n, sd in specs:
        xx = synth.field_data(f, sd, 1 << log_n)
        b = sfntt.DeviceBuffer(c, xx.size)
        b.upload(xx)
        bufs.append(b)
        xs.append(xx)
        descs.append(L.JobDesc(f, L.FORWARD, log_n, 1, b.ptr, b.ptr))
    arr = (L.JobDesc * len(descs))(*descs)
    plan = ctypes.c_void_p()
    c.check(c._lib.sfn_plan_create(c.ptr, arr, len(descs), 1, ctypes.byref(plan)))
    c.twiddles_limit(1)
    for k in range(3):
        c.coset_lde(x, 12, 8, 1, orc.bb_to_monty(101 + k))
    c.check(c._lib.sfn_plan_launch(plan, None))
    c.check(c._lib.sfn_synchronize(c.ptr, None))
    for (f, log_n, sd), b, xx in zip(specs, bufs, xs):
        assert np.array_equal(b.download(), orc.ntt_batch(f, False, log_n, 1, xx)), (f, log_n)
    c._lib.sfn_plan_destroy(plan)
    c.twiddles_limit(0)
    for b in bufs:
        b.free()
    c.close()


def test_server_fuses_jobs_of_every_ntt_shape_and_same_shape_ldes(ctx):
    # queued NTT jobs of different fields, sizes and directions run as one
    # many-job launch set; queued LDE jobs of one shape run as one batched LDE
    # (host memory: staged back to back; device memory: gathered and scattered)
    srv = make_server(ctx, inflight=1)
    keep = []
    for dev in (0, 1):
        for i in range(40):
            if i % 2:
                kind, f, log_n, batch, a, d = "lde", 1, 11, 1 + i % 5, 1, 0
            else:
                kind, f, log_n, batch, a, d = "ntt", (0, 1, 2)[i % 3], 8 + i % 9, 1, 0, (i // 2) % 2
            x = synth.field_data(f, 5000 + i + 100 * dev, batch << log_n)
            nout = batch << (log_n + a)
            if dev:
                din, dout = sfntt.DeviceBuffer(ctx, x.size), sfntt.DeviceBuffer(ctx, nout)
                din.upload(x)
                inp, outp, out = din.ptr, dout.ptr, dout
            else:
                out = np.zeros(nout, np.uint32)
                inp, outp, din = x.ctypes.data, out.ctypes.data, None
            r = L.Request(L.JOB_LDE if kind == "lde" else L.JOB_NTT, f, d, log_n, batch, a, 0, dev, inp, outp)
            t = ctypes.c_uint64()
            ctx.check(ctx._lib.sfn_server_submit(srv, ctypes.byref(r), ctypes.byref(t)))
            keep.append((t.value, kind, f, log_n, batch, a, d, x, out, dev, din))
    groups = []
    for t, kind, f, log_n, batch, a, d, x, out, dev, din in keep:
        tm = L.JobTiming()
        ctx.check(ctx._lib.sfn_server_wait(srv, t, ctypes.byref(tm)))
        groups.append(tm.group_size)
        got = out.download() if dev else out
        exp = orc.coset_lde_batch(x, log_n, batch, a) if kind == "lde" else orc.ntt_batch(f, d == 1, log_n, batch, x)
        assert np.array_equal(got, exp), (kind, f, log_n, batch, d, dev)
    ctx._lib.sfn_server_destroy(srv)
    assert max(groups) > 1


def test_lone_large_job_starts_while_workers_sleep(ctx):
    """A large job submitted while every worker sleeps starts at once.  The
    small-job lane cannot run it, so a wake-up that reached only that worker
    would leave the job queued until the next submit (regression test)."""
    log_n = 23  # 2^23 output elements: a large job
    x = synth.field_data(0, 77, 1 << log_n)
    din, dout = sfntt.DeviceBuffer(ctx, x.size), sfntt.DeviceBuffer(ctx, x.size)
    din.upload(x)
    lib = ctx._lib
    srv = make_server(ctx, inflight=4)
    try:
        for rep in range(12):
            time.sleep(0.01)  # idle workers poll for 100 us, then sleep
            r = L.Request(L.JOB_NTT, 0, L.FORWARD, log_n, 1, 0, 0, 1, din.ptr, dout.ptr)
            t = ctypes.c_uint64()
            ctx.check(lib.sfn_server_submit(srv, ctypes.byref(r), ctypes.byref(t)))
            tm = L.JobTiming()
            deadline = time.time() + 5.0
            while True:
                st = lib.sfn_server_poll(srv, t.value, ctypes.byref(tm))
                if st != L.ERR_NOT_READY:
                    break
                assert time.time() < deadline, f"repeat {rep}: the job never started"
                time.sleep(0.0005)
            ctx.check(st)
            assert tm.start_ns - tm.submit_ns < 50_000_000, f"repeat {rep}: started after {tm.start_ns - tm.submit_ns} ns"
    finally:
        lib.sfn_server_destroy(srv)
    assert np.array_equal(dout.download(), orc.m31_forward(x))
    din.free()
    dout.free()
