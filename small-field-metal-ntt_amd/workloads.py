"""Secondary benchmark workloads (bench.py --workload ...), one JSON dict each.

bb16    BabyBear forward NTT, one column of 2^16, every variant
        (benches/ntt_benchmark.rs:632-700 shootout shape, seed 12345 + n)
m31_20  M31 circle forward NTT, one column of 2^20, every variant, twiddles
        resident on the device (benches/ntt_benchmark.rs:183-375 shape)
        ("auto" is the autotuner's schedule for the shape, tuned before timing)
uneven  512 BabyBear/M31 columns with log_n ~ U{10..22} in one multi-job plan
        (one launch per pass, hipGraph replay) against launching every column
        on its own
small   many small transforms (2^8 and 2^10 points, both fields; the batched
        small-NTT regime of lattice schemes): one batched job, a plan of
        single-column jobs (hipGraph), and single-column launches
kb16    KoalaBear forward NTT, one column of 2^16, every variant (the bb16
        shape on the field the reference lists as planned, README.md:20)
batch20 forward NTT of 256 columns of 2^20 (the LDE's matrix shape) in every
        field, SoA and row-major, one in-place batched call each

Device time is a hipEvent pair around the launch sequence with inputs
resident in HBM; "call" time is the host-pointer entry point's wall clock
(upload + launches + download), the boundary the reference's shootout timed.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

import sfntt
from sfntt import _lib as L
from sfntt import synth

FIELD_NAMES = {L.FIELD_M31: "M31", L.FIELD_BABYBEAR: "BabyBear", L.FIELD_KOALABEAR: "KoalaBear"}
FIELD_TAGS = {L.FIELD_M31: "m31", L.FIELD_BABYBEAR: "bb", L.FIELD_KOALABEAR: "kb"}
VARIANTS = [("v1_ct_dit_r2", L.CT_DIT_R2), ("v2_ct_gs_r2", L.CT_GS_R2), ("v3_stockham_r2", L.STOCKHAM_R2),
            ("v4_ct_gs_r4", L.CT_GS_R4), ("auto", L.VARIANT_AUTO)]


def _single(ctx, field: int, log_n: int, steps: int, warmup: int) -> dict:
    lib = ctx._lib
    n = 1 << log_n
    x = synth.field_data(field, 12345 + n, n)
    din, dout = sfntt.DeviceBuffer(ctx, n), sfntt.DeviceBuffer(ctx, n)
    din.upload(x)
    ctx.twiddles_prepare(field, log_n)
    # AUTO runs the autotuner's schedule for this shape (tuning is not timed)
    best = ctypes.c_uint32()
    ctx.check(lib.sfn_autotune(ctx.ptr, field, log_n, 1, L.LAYOUT_SOA, ctypes.byref(best)))
    res = {}
    ns = ctypes.c_uint64()
    ctx.check(lib.sfn_call_timing(ctx.ptr, 1))  # device time per call (event pair)
    for name, v in VARIANTS:
        dev = []
        for k in range(warmup + steps):
            ctx.check(lib.sfn_ntt_device(ctx.ptr, field, v, L.FORWARD, log_n, 1, L.LAYOUT_SOA, din.ptr, dout.ptr,
                                         None))
            ctx.check(lib.sfn_last_device_ns(ctx.ptr, ctypes.byref(ns)))
            if k >= warmup:
                dev.append(ns.value / 1e3)
        wall = []
        for k in range(warmup + steps):
            t0 = time.perf_counter()
            ctx.ntt(field, v, L.FORWARD, log_n, 1, x)
            if k >= warmup:
                wall.append((time.perf_counter() - t0) * 1e6)
        d, w = float(np.median(dev)), float(np.median(wall))
        res[name] = {"device_us": round(d, 2), "call_us": round(w, 1),
                     "device_melem_s": round(n / d, 1), "call_melem_s": round(n / w, 1)}
    ctx.check(lib.sfn_call_timing(ctx.ptr, 0))
    return res


def run(name: str, ctx, steps: int, warmup: int) -> dict:
    if name in ("bb16", "m31_20", "kb16"):
        field, log_n = {"bb16": (L.FIELD_BABYBEAR, 16), "m31_20": (L.FIELD_M31, 20),
                        "kb16": (L.FIELD_KOALABEAR, 16)}[name]
        r = _single(ctx, field, log_n, max(steps, 20), max(warmup, 5))
        best = min(r.values(), key=lambda e: e["device_us"])
        return {"metric": f"{FIELD_NAMES[field]} forward NTT 2^{log_n} (one column), device Melem/s",
                "value": best["device_melem_s"], "unit": "Melem/s", "higher_is_better": True,
                "dtype": "u32", "data": "synthetic (reference LCG, seed 12345+n)",
                "config": {"workload": f"{FIELD_TAGS[field]}_forward_2^{log_n}_single_column",
                           "log_n": log_n, "batch": 1}, "variants": r}
    if name == "batch20":
        return _batch20(ctx, steps, warmup)
    if name == "uneven":
        return _uneven(ctx, steps, warmup)
    if name == "small":
        return _small(ctx, steps, warmup)
    raise ValueError(name)


def _batch20(ctx, steps, warmup):
    lib = ctx._lib
    log_n, cols = 20, 256
    total = cols << log_n
    buf = sfntt.DeviceBuffer(ctx, total)
    ns = ctypes.c_uint64()
    modes = {}
    ctx.check(lib.sfn_call_timing(ctx.ptr, 1))  # device time per call (event pair)
    for field in (L.FIELD_KOALABEAR, L.FIELD_BABYBEAR, L.FIELD_M31):
        buf.upload(synth.field_data(field, 42 + log_n, total))
        for lay, lname in ((L.LAYOUT_SOA, "soa"), (L.LAYOUT_ROW_MAJOR, "rowmajor")):
            dev = []
            for k in range(warmup + steps):
                ctx.check(lib.sfn_ntt_device(ctx.ptr, field, L.VARIANT_AUTO, L.FORWARD, log_n, cols, lay, buf.ptr,
                                             buf.ptr, None))
                ctx.check(lib.sfn_last_device_ns(ctx.ptr, ctypes.byref(ns)))
                if k >= warmup:
                    dev.append(ns.value / 1e3)
            d = float(np.median(dev))
            modes[f"{FIELD_TAGS[field]}_{lname}"] = {"device_us": round(d, 1),
                                                    "gelem_s": round(total / d / 1e3, 2),
                                                    "hbm_gb_s_min": round(2 * 4 * total / d / 1e3, 1)}
    ctx.check(lib.sfn_call_timing(ctx.ptr, 0))
    kb = modes["kb_rowmajor"]
    return {"metric": "KoalaBear forward NTT, 256 columns of 2^20 (row-major), device Gelem/s",
            "value": kb["gelem_s"], "unit": "Gelem/s", "higher_is_better": True, "dtype": "u32",
            "data": "synthetic (reference LCG, seed 42+log_n)",
            "config": {"workload": "forward_ntt_2^20x256_all_fields", "log_n": log_n, "batch": cols},
            "modes": modes}


def _timed(ctx, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    ctx.check(ctx._lib.sfn_synchronize(ctx.ptr, None))
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    ctx.check(ctx._lib.sfn_synchronize(ctx.ptr, None))
    return (time.perf_counter() - t0) / steps


def _small(ctx, steps, warmup):
    lib = ctx._lib
    out = {}
    for field, log_n, cols in ((L.FIELD_BABYBEAR, 8, 1 << 16), (L.FIELD_M31, 8, 1 << 16),
                               (L.FIELD_BABYBEAR, 10, 1 << 14), (L.FIELD_M31, 10, 1 << 14)):
        n = 1 << log_n
        total = cols * n
        buf = sfntt.DeviceBuffer(ctx, total)
        buf.upload(synth.field_data(field, 0x5A11 + log_n, total))
        key = f"{'bb' if field else 'm31'}_2^{log_n}_x{cols}"
        dt = _timed(ctx, lambda: ctx.check(lib.sfn_ntt_device(ctx.ptr, field, L.VARIANT_AUTO, L.FORWARD, log_n, cols,
                                                              L.LAYOUT_SOA, buf.ptr, buf.ptr, None)), steps, warmup)
        res = {"batch_job_us": round(dt * 1e6, 1), "batch_job_melem_s": round(total / dt / 1e6, 1)}
        descs = [L.JobDesc(field, L.FORWARD, log_n, 1, ctypes.c_void_p(buf.ptr.value + 4 * n * c),
                           ctypes.c_void_p(buf.ptr.value + 4 * n * c)) for c in range(cols)]
        arr = (L.JobDesc * cols)(*descs)
        for use_graph, tag in ((1, "plan_graph"), (0, "plan_direct")):
            plan = ctypes.c_void_p()
            ctx.check(lib.sfn_plan_create(ctx.ptr, arr, cols, use_graph, ctypes.byref(plan)))
            dt = _timed(ctx, lambda: ctx.check(lib.sfn_plan_launch(plan, None)), steps, warmup)
            lib.sfn_plan_destroy(plan)
            res.update({f"{tag}_us": round(dt * 1e6, 1), f"{tag}_melem_s": round(total / dt / 1e6, 1)})
        few = 256  # single-column launches, timed on a subset and scaled per column
        ptrs = [ctypes.c_void_p(buf.ptr.value + 4 * n * c) for c in range(few)]
        dt = _timed(ctx, lambda: [ctx.check(lib.sfn_ntt_device(ctx.ptr, field, L.VARIANT_AUTO, L.FORWARD, log_n, 1,
                                                                L.LAYOUT_SOA, p, p, None)) for p in ptrs],
                    max(1, steps // 5), 1)
        res.update({"per_column_launch_us": round(dt / few * 1e6, 2),
                    "per_column_melem_s": round(few * n / dt / 1e6, 1)})
        out[key] = res
        buf.free()
    best = max(max(v["plan_graph_melem_s"], v["plan_direct_melem_s"]) for v in out.values())
    return {"metric": "batched small NTTs (2^8, 2^10 points), Melem/s", "value": best, "unit": "Melem/s",
            "higher_is_better": True, "dtype": "u32", "data": "synthetic (reference LCG values)",
            "config": {"workload": "small_ntt_batches"}, "modes": out}


def _uneven(ctx, steps: int, warmup: int) -> dict:
    lib = ctx._lib
    cols = synth.uneven_columns()
    bufs, descs, total = [], [], 0
    for f, log_n, seed in cols:
        x = synth.field_data(f, seed, 1 << log_n)
        b = sfntt.DeviceBuffer(ctx, x.size)
        b.upload(x)
        bufs.append(b)
        descs.append(L.JobDesc(f, L.FORWARD, log_n, 1, b.ptr, b.ptr))
        total += 1 << log_n
    arr = (L.JobDesc * len(descs))(*descs)
    out = {}
    for use_graph in (1, 0):
        plan = ctypes.c_void_p()
        ctx.check(lib.sfn_plan_create(ctx.ptr, arr, len(descs), use_graph, ctypes.byref(plan)))
        nl, nt = ctypes.c_uint32(), ctypes.c_uint32()
        ctx.check(lib.sfn_plan_info(plan, ctypes.byref(nl), ctypes.byref(nt)))
        for _ in range(warmup):
            ctx.check(lib.sfn_plan_launch(plan, None))
        ctx.check(lib.sfn_synchronize(ctx.ptr, None))
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.check(lib.sfn_plan_launch(plan, None))
        ctx.check(lib.sfn_synchronize(ctx.ptr, None))
        dt = (time.perf_counter() - t0) / steps
        out["plan_graph" if use_graph else "plan_direct"] = {
            "ms": round(dt * 1e3, 3), "melem_s": round(total / dt / 1e6, 1), "launches": nl.value,
            "tiles": nt.value}
        lib.sfn_plan_destroy(plan)
    # one launch sequence per column (tuned kernels, same stream, no waits in between)
    for _ in range(warmup):
        for (f, log_n, _), b in zip(cols, bufs):
            ctx.check(lib.sfn_ntt_device(ctx.ptr, f, L.VARIANT_AUTO, L.FORWARD, log_n, 1, L.LAYOUT_SOA, b.ptr, b.ptr,
                                         None))
    ctx.check(lib.sfn_synchronize(ctx.ptr, None))
    t0 = time.perf_counter()
    for _ in range(steps):
        for (f, log_n, _), b in zip(cols, bufs):
            ctx.check(lib.sfn_ntt_device(ctx.ptr, f, L.VARIANT_AUTO, L.FORWARD, log_n, 1, L.LAYOUT_SOA, b.ptr, b.ptr,
                                         None))
    ctx.check(lib.sfn_synchronize(ctx.ptr, None))
    dt = (time.perf_counter() - t0) / steps
    out["per_column"] = {"ms": round(dt * 1e3, 3), "melem_s": round(total / dt / 1e6, 1), "launches": None}
    return {"metric": "uneven batch: 512 BabyBear+M31 columns, log_n ~ U{10..22}, forward, Melem/s",
            "value": max(out["plan_graph"]["melem_s"], out["plan_direct"]["melem_s"]), "unit": "Melem/s",
            "higher_is_better": True, "dtype": "u32",
            "data": "synthetic (SplitMix64 seed 0x5EED0004 shapes, reference LCG values)",
            "config": {"workload": "uneven_512_columns_log10to22", "columns": len(cols), "elements": total},
            "modes": out}
