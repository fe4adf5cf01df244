"""Runs the library's kernels and host runtime on the CPU model (see
build_model.py) and checks every result against the CPU oracle.

    LD_PRELOAD=$(g++ -print-file-name=libasan.so) python tests/cpu_model/run_model.py <suite> [...]

tests/test_cpu_model.py starts this in a subprocess with the ASan runtime
preloaded.  Each suite exits non-zero on the first mismatch, on any
buffer-resource access outside its resource, or on a sanitizer report.
Suites:
  ntt        every variant x field x direction, single columns and batches
             in both layouts, sizes across every pass split
  lde        coset LDE shapes of the GPU suite at small sizes, both layouts
             and shifts, every middle-pass width; natural row order
  schedules  every schedule the autotuner may force (NTT and LDE shapes)
  plans      uneven multi-job plans (direct and graph replays), contiguous
             small columns
  server     serving jobs alone and in a mix, host and device buffers
  faults     injected allocation / launch / copy failures: each call reports
             the error, releases what it held, and the next call succeeds
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "small-field-metal-ntt_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import oracle as orc  # noqa: E402
from sfntt import _lib as L  # noqa: E402

MODEL_LIB = os.path.join(HERE, "_build", "libsfntt_model.so")
L.LIB_PATH = MODEL_LIB  # the ctypes binding loads the model instead of the gfx950 library
import sfntt  # noqa: E402

lib = L.load()
lib.sfn_model_oob_count.restype = ctypes.c_uint64
lib.sfn_model_launch_count.restype = ctypes.c_uint64
lib.sfn_model_device_bytes.restype = ctypes.c_uint64
lib.sfn_model_peak_device_bytes.restype = ctypes.c_uint64
lib.sfn_model_malloc_count.restype = ctypes.c_uint64
lib.sfn_model_set_faults.argtypes = [ctypes.c_char_p]

M31, BB, KB = L.FIELD_M31, L.FIELD_BABYBEAR, L.FIELD_KOALABEAR
VARIANTS = [L.CT_DIT_R2, L.CT_GS_R2, L.STOCKHAM_R2, L.CT_GS_R4, L.VARIANT_AUTO]


def check(cond, what):
    if not cond:
        raise AssertionError(what)


def same(got, exp, what):
    got, exp = np.asarray(got), np.asarray(exp)
    if got.shape != exp.shape or not np.array_equal(got, exp):
        bad = np.flatnonzero(got != exp) if got.shape == exp.shape else [-1]
        raise AssertionError(f"{what}: {len(bad)} mismatches, first at {bad[0]}")
    oob = lib.sfn_model_oob_count()
    check(oob == 0, f"{what}: {oob} buffer accesses outside their resource")


def log(msg):
    print(msg, flush=True)


def to_rm(x, log_n, batch):
    return x.reshape(batch, 1 << log_n).T.copy().ravel()


def from_rm(y, log_n, batch):
    return y.reshape(1 << log_n, batch).T.ravel()


# ---------------------------------------------------------------- suites
def suite_ntt(ctx):
    for field in (M31, BB, KB):
        ref = {False: lambda x, f=field: orc.ntt_batch(f, False, len(x).bit_length() - 1, 1, x),
               True: lambda x, f=field: orc.ntt_batch(f, True, len(x).bit_length() - 1, 1, x)}
        for variant in VARIANTS:
            for log_n in (1, 2, 3, 5, 6, 7, 9, 11, 12, 13, 15):
                x = orc.lcg_fill(log_n * 7919 + variant, 1 << log_n, field)
                for d in (0, 1):
                    log(f"[ntt] field={field} variant={variant} log_n={log_n} dir={d}")
                    y, _ = ctx.ntt(field, variant, d, log_n, 1, x)
                    same(y, ref[d == 1](x), f"ntt field={field} v={variant} 2^{log_n} dir={d}")
    for field in (M31, BB, KB):
        for log_n, batch in ((4, 3), (8, 16), (10, 5), (12, 16), (14, 4), (16, 2)):
            x = orc.lcg_fill(99 + log_n, batch << log_n, field)
            for d in (0, 1):
                exp = orc.ntt_batch(field, d == 1, log_n, batch, x)
                for variant in (L.CT_GS_R4, L.VARIANT_AUTO):
                    log(f"[ntt] batch field={field} 2^{log_n}x{batch} dir={d} v={variant}")
                    y, _ = ctx.ntt(field, variant, d, log_n, batch, x)
                    same(y, exp, f"batch soa field={field} 2^{log_n}x{batch} dir={d} v={variant}")
                    y, _ = ctx.ntt(field, variant, d, log_n, batch, to_rm(x, log_n, batch), L.LAYOUT_ROW_MAJOR)
                    same(from_rm(y, log_n, batch), exp, f"batch row-major field={field} 2^{log_n}x{batch} dir={d}")


LDE_CASES = [(8, 4, 1), (12, 8, 1), (8, 4, 2), (1, 3, 1), (5, 2, 3), (10, 6, 2), (13, 3, 1), (11, 16, 1),
             (12, 32, 1), (14, 64, 1), (16, 16, 1), (10, 4, 0), (0, 3, 2)]


def suite_lde(ctx):
    for log_n, batch, a in LDE_CASES:
        x = orc.lcg_fill(42, batch << log_n, BB)
        for shift in (0, orc.bb_to_monty(31)):
            exp = orc.coset_lde_batch(x, log_n, batch, a, shift)
            log(f"[lde] 2^{log_n}x{batch} a={a} shift={shift}")
            y, _ = ctx.coset_lde(x, log_n, batch, a, shift, L.LAYOUT_SOA)
            same(y, exp, f"lde soa 2^{log_n}x{batch} a={a}")
            y, _ = ctx.coset_lde(to_rm(x, log_n, batch), log_n, batch, a, shift, L.LAYOUT_ROW_MAJOR)
            same(from_rm(y, log_n + a, batch), exp, f"lde row-major 2^{log_n}x{batch} a={a}")
    # host-pointer calls of more than one 32 MiB chunk run as column chunks on
    # several streams (SoA: contiguous; row-major: 2D copies of column windows)
    for log_n, batch, a in ((13, 2200, 1),):
        x = orc.lcg_fill(77 + log_n, batch << log_n, BB)
        exp = orc.coset_lde_batch(x, log_n, batch, a)
        log(f"[lde] chunked host call 2^{log_n}x{batch} a={a}")
        y, _ = ctx.coset_lde(x, log_n, batch, a)
        same(y, exp, f"chunked lde soa 2^{log_n}x{batch}")
        y, _ = ctx.coset_lde(to_rm(x, log_n, batch), log_n, batch, a, 0, L.LAYOUT_ROW_MAJOR)
        same(from_rm(y, log_n + a, batch), exp, f"chunked lde row-major 2^{log_n}x{batch}")
    x = orc.lcg_fill(5, 2200 << 13, KB)
    log("[lde] chunked host NTT 2^13x2200 inverse")
    y, _ = ctx.ntt(KB, L.VARIANT_AUTO, 1, 13, 2200, to_rm(x, 13, 2200), L.LAYOUT_ROW_MAJOR)
    same(from_rm(y, 13, 2200), orc.ntt_batch(KB, True, 13, 2200, x), "chunked ntt")
    import closed_forms as cf
    for log_n, batch, a in ((6, 5, 2), (12, 8, 1), (13, 16, 1), (10, 4, 0)):
        x = orc.lcg_fill(4242 + log_n, batch << log_n, BB)
        for layout in (L.LAYOUT_SOA, L.LAYOUT_ROW_MAJOR):
            rm = layout == L.LAYOUT_ROW_MAJOR
            shift = orc.bb_to_monty(31)
            exp = orc.coset_lde_batch(cf.rows_bitrev(x, log_n, batch, rm), log_n, batch, a, shift, layout=layout)
            exp = cf.rows_bitrev(exp, log_n + a, batch, rm)
            log(f"[lde] natural 2^{log_n}x{batch} a={a} layout={layout}")
            y, _ = ctx.coset_lde(x, log_n, batch, a, shift, layout, natural=True)
            same(y, exp, f"natural lde 2^{log_n}x{batch} a={a} layout={layout}")


def _autotune_candidates(log_n, lde):
    for mlt in range(6, 12):
        if (log_n + mlt - 1) // mlt > 3 or (lde and log_n > 2 * mlt):
            continue
        for logr in (3, 4, 5):
            for lw in (2, 3, 4, 5):
                for mlw in ((2, 3, 4, 5) if lde else (lw,)):
                    if mlt + max(lw, mlw) <= 15:
                        yield mlt, logr, lw, mlw


def _force(ctx, key_field, log_n, batch, layout, sched):
    mlt, logr, lw, mlw = sched
    js = (f'[{{"field":{key_field},"log_n":{log_n},"batch":{batch},"layout":{layout},'
          f'"max_log_t":{mlt},"logr":{logr},"log_w":{lw},"mid_log_w":{mlw}}}]')
    ctx.check(lib.sfn_tuning_load(ctx.ptr, js.encode()))


NOT_LAUNCHABLE = (sfntt.api.Unsupported, sfntt.api.InvalidArgument, sfntt.api.InvalidSize)


def suite_schedules(ctx, full=False):
    ntt_shapes = ((BB, 16, 4, L.LAYOUT_SOA), (BB, 16, 4, L.LAYOUT_ROW_MAJOR), (M31, 17, 3, L.LAYOUT_SOA),
                  (KB, 13, 3, L.LAYOUT_ROW_MAJOR))
    lde_shapes = ((12, 6, L.LAYOUT_SOA), (13, 8, L.LAYOUT_ROW_MAJOR))
    if not full:
        ntt_shapes, lde_shapes = ntt_shapes[3:], lde_shapes[1:]
    for field, log_n, batch, layout in ntt_shapes:
        x = orc.lcg_fill(1000 + log_n, batch << log_n, field)
        exp = [orc.ntt_batch(field, d == 1, log_n, batch, x, layout=layout) for d in (0, 1)]
        ran = 0
        for sched in _autotune_candidates(log_n, False):
            _force(ctx, field, log_n, batch, layout, sched)
            for d in (0, 1):
                log(f"[sched] field={field} 2^{log_n}x{batch} layout={layout} {sched} dir={d}")
                try:
                    y, _ = ctx.ntt(field, L.VARIANT_AUTO, d, log_n, batch, x, layout)
                except NOT_LAUNCHABLE:
                    continue
                same(y, exp[d], f"schedule {sched} field={field} dir={d}")
                ran += 1
        check(ran >= 20, f"only {ran} schedules ran")
    for log_n, batch, layout in lde_shapes:
        x = orc.lcg_fill(42 + log_n, batch << log_n, BB)
        shift = 0 if layout == L.LAYOUT_SOA else orc.bb_to_monty(31)
        exp = orc.coset_lde_batch(x, log_n, batch, 1, shift, layout)
        for sched in _autotune_candidates(log_n, True):
            log(f"[sched] lde 2^{log_n}x{batch} layout={layout} {sched}")
            _force(ctx, L.TUNE_LDE, log_n, batch, layout, sched)
            try:
                y, _ = ctx.coset_lde(x, log_n, batch, 1, shift, layout)
            except NOT_LAUNCHABLE:
                continue
            same(y, exp, f"lde schedule {sched}")


def _uneven(seed, count, lo, hi):
    st = orc.splitmix64_stream(seed)
    out = []
    for _ in range(count):
        f = next(st) & 1
        log_n = lo + next(st) % (hi - lo + 1)
        out.append((int(f), int(log_n), int(next(st))))
    return out


def run_plan(ctx, specs, use_graph, dirs, reps=1):
    from sfntt import synth
    bufs, descs, xs = [], [], []
    for k, (f, log_n, s) in enumerate(specs):
        x = synth.field_data(f, s, 1 << log_n)
        b = sfntt.DeviceBuffer(ctx, x.size)
        b.upload(x)
        bufs.append(b)
        xs.append(x)
        descs.append(L.JobDesc(f, dirs[k], log_n, 1, b.ptr, b.ptr))
    arr = (L.JobDesc * len(descs))(*descs)
    plan = ctypes.c_void_p()
    ctx.check(lib.sfn_plan_create(ctx.ptr, arr, len(descs), int(use_graph), ctypes.byref(plan)))
    for _ in range(reps):
        ctx.check(lib.sfn_plan_launch(plan, None))
    ctx.check(lib.sfn_synchronize(ctx.ptr, None))
    outs = [b.download() for b in bufs]
    lib.sfn_plan_destroy(plan)
    for b in bufs:
        b.free()
    return xs, outs


def suite_plans(ctx):
    for use_graph in (0, 1):
        specs = _uneven(0x5EED0004, 24, 6, 16) + [(2, 12, 5), (2, 9, 6)]
        dirs = [k % 2 for k in range(len(specs))]
        log(f"[plans] uneven graph={use_graph}")
        xs, outs = run_plan(ctx, specs, use_graph, dirs, reps=1)
        for (f, log_n, s), d, x, y in zip(specs, dirs, xs, outs):
            same(y, orc.ntt_batch(f, d == 1, log_n, 1, x), f"plan job field={f} 2^{log_n} dir={d}")
    # 22-layer columns (three passes) beside short ones
    specs = [(1, 22, 5), (0, 22, 6), (1, 10, 7), (0, 11, 8)]
    log("[plans] 2^22 columns")
    xs, outs = run_plan(ctx, specs, 1, [0, 1, 1, 0])
    for (f, log_n, s), d, x, y in zip(specs, [0, 1, 1, 0], xs, outs):
        same(y, orc.ntt_batch(f, d == 1, log_n, 1, x), f"plan 2^22 job field={f} 2^{log_n} dir={d}")


def _serve(ctx, reqs, device_ptrs, inflight=4, mem=0, faults=None, group=0):
    """Submits every request, then waits for each; returns (status, timing, expected, got) per job."""
    from sfntt import synth
    cfg = L.ServerConfig(mem, inflight, 0, group, 0)
    srv = ctypes.c_void_p()
    ctx.check(lib.sfn_server_create(ctx.ptr, ctypes.byref(cfg), ctypes.byref(srv)))
    if faults:
        lib.sfn_model_set_faults(faults.encode())
    keep = []
    for i, (kind, f, log_n, batch, a) in enumerate(reqs):
        x = synth.field_data(f, 1000 + i, batch << log_n)
        nout = batch << (log_n + a)
        if device_ptrs:
            din, dout = sfntt.DeviceBuffer(ctx, x.size), sfntt.DeviceBuffer(ctx, nout)
            din.upload(x)
            inp, outp, out = din.ptr, dout.ptr, dout
        else:
            out = np.zeros(nout, np.uint32)
            inp, outp, din = x.ctypes.data, out.ctypes.data, None
        d = i % 2 if kind == "ntt" else 0
        r = L.Request(L.JOB_LDE if kind == "lde" else L.JOB_NTT, f, d, log_n, batch, a, 0, device_ptrs, inp, outp)
        t = ctypes.c_uint64()
        ctx.check(lib.sfn_server_submit(srv, ctypes.byref(r), ctypes.byref(t)))
        keep.append((t.value, kind, f, log_n, batch, a, d, x, out, din))
    results = []
    for t, kind, f, log_n, batch, a, d, x, out, din in keep:
        tm = L.JobTiming()
        st = lib.sfn_server_wait(srv, t, ctypes.byref(tm))
        got = out.download() if device_ptrs else out
        results.append((st, tm.group_size, (kind, f, log_n, batch, a, d), _expect(kind, f, log_n, batch, a, d, x), got))
        if din is not None:
            din.free()
            out.free()
    lib.sfn_model_set_faults(b"")
    lib.sfn_server_destroy(srv)
    return results


def _held(ctx):
    """Device bytes outside the context's tables (server staging, descriptors, buffers)."""
    return lib.sfn_model_device_bytes() - ctx.twiddles_bytes()


def _expect(kind, f, log_n, batch, a, d, x):
    if kind == "lde":
        return orc.coset_lde_batch(x, log_n, batch, a)
    return orc.ntt_batch(f, d == 1, log_n, batch, x)


def suite_server(ctx):
    rng = np.random.default_rng(5)
    reqs = []
    for _ in range(40):
        if rng.random() < 0.6:
            reqs.append(("lde", BB, int(rng.integers(6, 12)), int(rng.integers(1, 9)), int(rng.integers(1, 3))))
        else:
            reqs.append(("ntt", int(rng.integers(0, 3)), int(rng.integers(8, 14)), 1, 0))
    # same-shape LDE jobs (fused into one batched LDE) and NTT jobs of every shape (one many-job launch)
    reqs += [("lde", BB, 9, 3, 1)] * 12 + [("ntt", int(rng.integers(0, 3)), int(rng.integers(6, 15)), 1, 0)
                                          for _ in range(24)]
    for device_ptrs in (0, 1):
        log(f"[server] {len(reqs)} mixed jobs device_ptrs={device_ptrs}")
        res = _serve(ctx, reqs, device_ptrs)
        for st, g, shape, exp, got in res:
            check(st == 0, f"job {shape} status {st}")
            same(got, exp, f"served {shape} device_ptrs={device_ptrs}")
        groups = [g for *_, g in [(r[0], r[1]) for r in res]]
        check(max(groups) > 1, "no job was fused")
        log(f"[server] largest group {max(groups)}, mean {np.mean(groups):.2f}")
    # a large host-memory job runs in column chunks on the worker's transfer streams
    log("[server] large host-memory LDE (column chunks)")
    for st, g, shape, exp, got in _serve(ctx, [("lde", BB, 14, 1100, 1), ("ntt", M31, 10, 1, 0)], 0, inflight=2):
        check(st == 0, f"job status {st}")
        same(got, exp, f"large host job {shape}")
    # admission by footprint: host jobs of 1 MiB in + 2 MiB out under a 7 MiB budget
    base = _held(ctx)
    lib.sfn_model_reset_counters()
    reqs = [("lde", BB, 12, 64, 1)] * 6
    log("[server] 6 LDE jobs of 3 MiB staging under a 7 MiB budget")
    for st, g, shape, exp, got in _serve(ctx, reqs, 0, inflight=3, mem=7 << 20):
        check(st == 0, f"job status {st}")
        same(got, exp, "LDE under a small budget")
    peak = lib.sfn_model_peak_device_bytes() - base
    log(f"[server] peak device bytes above the start {peak}")
    check(peak <= (7 << 20) + (1 << 20), f"staging peak {peak} exceeds the budget")
    check(_held(ctx) == base, "the server kept device memory after destroy")


def suite_server_failures(ctx):
    """An injected launch failure fails its group's jobs with GPU_EXECUTION;
    the other jobs complete, and the failed worker's staging is released."""
    base = _held(ctx)
    reqs = [("ntt", M31, 10 + k % 3, 1, 0) for k in range(12)] + [("lde", BB, 8, 2, 1)] * 6
    for device_ptrs in (0, 1):
        # host-memory groups run one at a time and fuse what queued meanwhile,
        # so the 18 jobs may take as few as five launches
        for n in (1, 2, 4):
            log(f"[server] launch failure {n} device_ptrs={device_ptrs}")
            res = _serve(ctx, reqs, device_ptrs, inflight=2, faults=f"launch:{n}")
            failed = [r for r in res if r[0] != 0]
            check(failed, "the injected failure reached no job")
            check(all(r[0] == L.ERR_GPU_EXECUTION for r in failed), "failed jobs report GPU_EXECUTION")
            for st, g, shape, exp, got in res:
                if st == 0:
                    same(got, exp, f"job {shape} beside a failure")
    # staging allocation failure: the job reports BUFFER_ALLOC_FAILED, later jobs run
    log("[server] staging allocation failure")
    res = _serve(ctx, [("lde", BB, 10, 4, 1)] * 3, 0, inflight=1, faults="malloc:1", group=1)
    check(res[0][0] == L.ERR_BUFFER_ALLOC_FAILED, f"first job status {res[0][0]}")
    for st, g, shape, exp, got in res[1:]:
        check(st == 0, f"status {st} after the failed allocation")
        same(got, exp, "LDE after a failed staging allocation")
    check(_held(ctx) == base, "device memory held after the failures")
    # destroy with jobs still queued: waiters return (the queued ones with BUSY), nothing hangs
    import threading
    from sfntt import synth
    cfg = L.ServerConfig(0, 1, 0, 1, 0)
    srv = ctypes.c_void_p()
    ctx.check(lib.sfn_server_create(ctx.ptr, ctypes.byref(cfg), ctypes.byref(srv)))
    x = synth.field_data(1, 5, 8 << 10)
    outs, tickets = [], []
    for i in range(20):
        out = np.zeros(16 << 10, np.uint32)
        r = L.Request(L.JOB_LDE, BB, 0, 10, 8, 1, 0, 0, x.ctypes.data, out.ctypes.data)
        t = ctypes.c_uint64()
        ctx.check(lib.sfn_server_submit(srv, ctypes.byref(r), ctypes.byref(t)))
        outs.append(out)
        tickets.append(t.value)
    statuses = {}

    def waiter(t):
        statuses[t] = lib.sfn_server_wait(srv, t, None)

    ths = [threading.Thread(target=waiter, args=(t,)) for t in tickets[-5:]]
    for th in ths:
        th.start()
    log("[server] destroy with queued jobs")
    lib.sfn_server_destroy(srv)
    for th in ths:
        th.join(timeout=60)
        check(not th.is_alive(), "a waiter is still blocked after destroy")
    check(all(v in (0, L.ERR_BUSY) for v in statuses.values()), f"waiter statuses {statuses}")
    check(_held(ctx) == base, "device memory held after destroy")


def suite_cache(ctx):
    """The per-shape tables stay within the byte budget across many distinct
    LDE shifts, and a plan's tables survive evictions until it is destroyed."""
    c = sfntt.HipContext(0)
    x = orc.lcg_fill(3, 4 << 10, BB)
    y, _ = c.coset_lde(x, 10, 4, 1, orc.bb_to_monty(3))
    base = c.twiddles_bytes()
    limit = 64 << 10
    c.check(lib.sfn_twiddles_limit(c.ptr, limit))
    for k in range(40):
        shift = orc.bb_to_monty(3 + 2 * k)
        log(f"[cache] LDE with shift {3 + 2 * k}")
        y, _ = c.coset_lde(x, 10, 4, 1, shift)
        same(y, orc.coset_lde_batch(x, 10, 4, 1, shift), f"LDE shift {3 + 2 * k}")
        check(c.twiddles_bytes() <= base + limit, f"tables {c.twiddles_bytes()} past the budget")
    # a plan pins its tables: with the budget at one byte everything else is evicted
    specs = [(0, 13, 11), (1, 14, 12), (0, 12, 13), (1, 15, 14)]
    from sfntt import synth
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
    c.check(lib.sfn_plan_create(c.ptr, arr, len(descs), 1, ctypes.byref(plan)))
    c.check(lib.sfn_twiddles_limit(c.ptr, 1))
    for k in range(3):  # other calls evict and rebuild their own tables meanwhile
        c.coset_lde(x, 10, 4, 1, orc.bb_to_monty(101 + k))
    log("[cache] plan launch after evictions")
    c.check(lib.sfn_plan_launch(plan, None))
    c.check(lib.sfn_synchronize(c.ptr, None))
    for (f, log_n, sd), b, xx in zip(specs, bufs, xs):
        same(b.download(), orc.ntt_batch(f, False, log_n, 1, xx), f"plan job {f} 2^{log_n} after evictions")
    lib.sfn_plan_destroy(plan)
    c.check(lib.sfn_twiddles_limit(c.ptr, 0))
    for b in bufs:
        b.free()
    c.close()


def suite_faults(ctx):
    """Every injected failure surfaces as a status, and the next call works."""
    x = orc.lcg_fill(7, 8 << 10, BB)
    exp_lde = orc.coset_lde_batch(x, 10, 8, 1)
    exp_ntt = orc.ntt_batch(BB, False, 10, 8, x)
    kinds = {"malloc": sfntt.api.BufferAllocFailed, "launch": sfntt.api.GpuExecutionError,
             "memcpy": sfntt.api.GpuExecutionError}
    for what, exc in kinds.items():
        for n in (1, 2, 3):
            c = sfntt.HipContext(0)
            lib.sfn_model_set_faults(f"{what}:{n}".encode())
            log(f"[faults] {what} call {n}: host-pointer LDE")
            try:
                y, _ = c.coset_lde(x, 10, 8, 1)
                same(y, exp_lde, f"LDE with {what}:{n} that did not trigger")
            except exc:
                pass
            lib.sfn_model_set_faults(b"")
            y, _ = c.coset_lde(x, 10, 8, 1)
            same(y, exp_lde, f"LDE after an injected {what} failure")
            y, _ = c.ntt(BB, L.VARIANT_AUTO, 0, 10, 8, x)
            same(y, exp_ntt, f"NTT after an injected {what} failure")
            c.close()
    check(lib.sfn_model_device_bytes() == 0, f"{lib.sfn_model_device_bytes()} device bytes still held")


def suite_noalloc(ctx):
    """The allocation-free natural-order LDE makes no device allocation once the
    shape's tables are resident, gives the allocating entry point's bits, and
    refuses shapes whose tables are missing or whose output exceeds its scratch."""
    import closed_forms as cf
    c = sfntt.HipContext(0)
    fn, ref = lib.sfn_coset_lde_natural_device_noalloc, lib.sfn_coset_lde_natural_device
    for log_n, batch, a, layout in ((13, 16, 1, L.LAYOUT_ROW_MAJOR), (10, 8, 2, L.LAYOUT_ROW_MAJOR),
                                    (12, 4, 1, L.LAYOUT_SOA)):
        rm = layout == L.LAYOUT_ROW_MAJOR
        shift = orc.bb_to_monty(31)
        x = orc.lcg_fill(808 + log_n, batch << log_n, BB)
        nout = batch << (log_n + a)
        din, dref, dna = sfntt.DeviceBuffer(c, x.size), sfntt.DeviceBuffer(c, nout), sfntt.DeviceBuffer(c, nout)
        din.upload(x)
        log(f"[noalloc] 2^{log_n}x{batch} a={a} layout={layout}")
        c.check(ref(c.ptr, log_n, batch, a, shift, layout, din.ptr, dref.ptr, None))
        before = lib.sfn_model_malloc_count()
        c.check(fn(c.ptr, log_n, batch, a, shift, layout, din.ptr, dna.ptr, None))
        check(lib.sfn_model_malloc_count() == before, "the allocation-free call allocated")
        exp = orc.coset_lde_batch(cf.rows_bitrev(x, log_n, batch, rm), log_n, batch, a, shift, layout=layout)
        exp = cf.rows_bitrev(exp, log_n + a, batch, rm)
        same(dna.download(), dref.download(), f"noalloc vs allocating 2^{log_n}x{batch}")
        same(dna.download(), exp, f"noalloc vs oracle 2^{log_n}x{batch}")
        for b in (din, dref, dna):
            b.free()
    d = sfntt.DeviceBuffer(c, 3 << 20)
    launches = lib.sfn_model_launch_count()
    rc = fn(c.ptr, 11, 8, 1, orc.bb_to_monty(31), L.LAYOUT_ROW_MAJOR, d.ptr,
            ctypes.c_void_p(d.ptr.value + (8 << 13)), None)
    check(rc == L.ERR_NOT_READY, f"missing tables: status {rc}")
    check(lib.sfn_model_launch_count() == launches, "a refused call launched kernels")
    rc = fn(c.ptr, 20, 16, 1, 0, L.LAYOUT_ROW_MAJOR, d.ptr, ctypes.c_void_p(d.ptr.value + 4), None)
    check(rc == L.ERR_INVALID_ARGUMENT, f"overlap: status {rc}")
    big = sfntt.DeviceBuffer(c, 48 << 20)
    rc = fn(c.ptr, 20, 16, 1, 0, L.LAYOUT_ROW_MAJOR, big.ptr, ctypes.c_void_p(big.ptr.value + (64 << 20)), None)
    check(rc == L.ERR_INVALID_SIZE, f"over the scratch: status {rc}")
    d.free()
    big.free()
    c.close()


SUITES = {"ntt": suite_ntt, "lde": suite_lde, "schedules": suite_schedules,
          "schedules_full": lambda ctx: suite_schedules(ctx, full=True), "plans": suite_plans,
          "server": lambda ctx: (suite_server(ctx), suite_server_failures(ctx)), "faults": suite_faults,
          "cache": suite_cache, "noalloc": suite_noalloc}


def main(argv):
    names = argv or list(SUITES)
    for name in names:
        lib.sfn_model_reset_counters()
        ctx = sfntt.HipContext(0)
        SUITES[name](ctx)
        ctx.close()
        log(f"[model] suite {name}: ok ({lib.sfn_model_launch_count()} kernel launches, 0 out-of-range accesses)")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
