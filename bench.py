#!/usr/bin/env python3
"""Headline benchmark: BabyBear batched coset LDE, 2^20 rows x 256 columns, 2x
blowup (2^21 rows out), on one MI355X.

A step is one LDE over the whole batch (iDFT -> n^-1 * shift^j -> zero-pad ->
DFT) with the input already resident in HBM.  `value` is output elements per
second (Melem/s), the reference's LDE throughput unit (benches/ntt_benchmark.rs:516-517).
The default workload is the configuration as stated (row-major batch layout,
shift = field generator 31); --semantics reference runs the reference's own
SoA layout with shift = omega_{2N} (src/lde/mod.rs:59).  Both produce
bit-exact results checked by the tests.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "small-field-metal-ntt_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
COUNTERS_FILE = os.path.join(ROOT, "profiles", "lde_2e20x256_hbm_counters.txt")
VALU_FILE = os.path.join(ROOT, "profiles", "lde_2e20x256_valu_counters.txt")
# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per 4 cycles per SIMD, 2.4 GHz
VALU_PEAK_WAVE_INSTS = 256 * 4 * 2.4e9 / 4


def kernel_bytes(label: str, log_n: int, batch: int, added_bits: int) -> int:
    """Algorithmic bytes of one launch: each value it must read once + each it writes once."""
    n, N = batch << log_n, batch << (log_n + added_bits)
    if "rows_bitrev" in label:  # row-order permutation: every row read and written once
        return 8 * (n if "<in>" in label else N)
    if "mid" in label.split("<")[0]:  # fused middle pass: n in, N out
        return 4 * (n + N)
    if "inv" in label:  # iDFT low layers: n in, n out
        return 8 * n
    return 8 * N        # DFT low layers over the extended columns


def traffic_key(label: str) -> str:
    """Kernel label (sfn_pass_label) -> key of the counters summary:
    k_mid_fast, k_fast_inv, k_fast_fwd, ..."""
    head = label.split("<")[0]
    if head == "k_fast" or head == "k_pass":
        return head + ("_inv" if ",inv," in label else "_fwd")
    return head


def read_valu(key: str):
    """VALU (wave64) instructions per launch from the committed rocprofv3 --pmc summary, if present."""
    if not os.path.exists(VALU_FILE):
        return None
    for line in open(VALU_FILE):
        parts = line.split("#")[0].split()
        if parts and parts[0] == key:
            kv = dict(p.split("=", 1) for p in parts[1:] if "=" in p)
            if "valu_insts" in kv:
                return float(kv["valu_insts"])
    return None


def read_traffic(key: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    if not os.path.exists(COUNTERS_FILE):
        return None
    for line in open(COUNTERS_FILE):
        parts = line.split()
        if parts and parts[0] == key and len(parts) >= 2:
            kv = dict(p.split("=", 1) for p in parts[1:] if "=" in p)
            if "hbm_bytes" in kv:
                return float(kv["hbm_bytes"])
    return None


def host_cpu() -> str:
    """nproc and the CPU model of this host (recorded beside the CPU baselines)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return f"{os.cpu_count()} x {model}"


def cpu_baseline(x_soa_sample, log_n: int, cols: int, added_bits: int, shift_raw: int,
                 gpu_first_cols=None) -> dict:
    """The oracle's restatement of coset_lde_batch_cpu (src/lde/mod.rs:127-176),
    one thread, on the first `cols` columns of the same input."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as orc

    orc.lib()
    t0 = time.perf_counter()
    out = orc.coset_lde_batch(x_soa_sample, log_n, cols, added_bits, shift_raw, threads=1)
    dt = time.perf_counter() - t0
    res = {"value": round(cols * (1 << (log_n + added_bits)) / dt / 1e6, 3), "unit": "Melem/s",
           "cores": 1, "kind": "port",
           "sample": f"{cols} of 256 columns, 2^{log_n} rows, {1 << added_bits}x, SoA, {dt:.2f} s, "
                     f"C restatement of coset_lde_batch_cpu, 1 thread", "host": host_cpu()}
    if gpu_first_cols is not None:
        res["parity_sample"] = bool(np.array_equal(out, gpu_first_cols))
    return res


def cpu_baseline_all_cores(x_soa_sample, log_n: int, cols: int, added_bits: int, shift_raw: int) -> dict:
    """Same restatement with the columns split over every host thread (at most 16)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    threads = max(1, min(16, os.cpu_count() or 1))
    orc.lib()
    t0 = time.perf_counter()
    orc.coset_lde_batch(x_soa_sample, log_n, cols, added_bits, shift_raw, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(cols * (1 << (log_n + added_bits)) / dt / 1e6, 3), "unit": "Melem/s",
            "cores": threads, "kind": "port",
            "sample": f"{cols} of 256 columns, {dt:.2f} s, columns split over {threads} threads",
            "host": host_cpu()}


def end_to_end(ctx, x_dev, log_n: int, B: int, a: int, shift_raw: int, layout: int, steps: int) -> dict:
    """Host-pointer entry point (upload -> LDE -> download) on pinned host buffers."""
    import numpy as np

    lib = ctx._lib
    n_in, n_out = x_dev.size, B << (log_n + a)
    hin, hout = ctypes.c_void_p(), ctypes.c_void_p()
    ctx.check(lib.sfn_host_alloc(ctx.ptr, n_in * 4, ctypes.byref(hin)))
    ctx.check(lib.sfn_host_alloc(ctx.ptr, n_out * 4, ctypes.byref(hout)))
    try:
        ctypes.memmove(hin, x_dev.ctypes.data, n_in * 4)
        dev_ns = ctypes.c_uint64()
        pin = ctypes.cast(hin, ctypes.POINTER(ctypes.c_uint32))
        pout = ctypes.cast(hout, ctypes.POINTER(ctypes.c_uint32))
        times = []
        for k in range(steps + 1):
            t0 = time.perf_counter()
            ctx.check(lib.sfn_coset_lde_batch(ctx.ptr, log_n, B, a, shift_raw, layout, pin, pout,
                                              ctypes.byref(dev_ns)))
            if k:
                times.append(time.perf_counter() - t0)
        dt = float(np.median(times))
        return {"ms": round(dt * 1e3, 3), "melem_s": round(n_out / dt / 1e6, 2),
                "pcie_bytes": 4 * (n_in + n_out), "host_buffers": "pinned (sfn_host_alloc)",
                "steps": steps}
    finally:
        lib.sfn_host_free(ctx.ptr, hin)
        lib.sfn_host_free(ctx.ptr, hout)


def cpu_baseline_ntt(workload: str) -> dict:
    """The oracle's restatement of the reference CPU forward NTT (1 thread), 8
    columns in one call so the twiddle table is built once (the reference's
    TwiddleCache keeps it between calls)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as orc
    from sfntt import synth

    field, log_n = {"bb16": (1, 16), "m31_20": (0, 20), "kb16": (2, 16)}[workload]
    n, cols = 1 << log_n, 8
    x = np.concatenate([synth.field_data(field, 12345 + n, n)] * cols)
    orc.lib()
    t0 = time.perf_counter()
    orc.ntt_batch(field, False, log_n, cols, x, threads=1)
    dt = (time.perf_counter() - t0) / cols
    return {"value": round(n / dt / 1e6, 3), "unit": "Melem/s", "cores": 1, "kind": "port",
            "sample": f"{cols} columns of 2^{log_n} in one call, {dt * 1e6:.0f} us per column, 1 thread",
            "host": host_cpu()}


def dispersion(samples_ms) -> dict:
    """Median, min, max, mean, stddev, p95 and coefficient of variation of
    per-step times (benches/ntt_benchmark.rs:75-120 reports the same set)."""
    import numpy as np

    a = np.asarray(samples_ms, dtype=np.float64)
    mean = float(a.mean())
    sd = float(a.std(ddof=1)) if a.size > 1 else 0.0
    cv = 100.0 * sd / mean if mean > 0 else 0.0
    return {"median_ms": round(float(np.median(a)), 4), "min_ms": round(float(a.min()), 4),
            "max_ms": round(float(a.max()), 4), "mean_ms": round(mean, 4), "stddev_ms": round(sd, 4),
            "p95_ms": round(float(np.percentile(a, 95)), 4), "cv_pct": round(cv, 2),
            "high_variance": bool(cv > 10.0), "samples": int(a.size)}


def step_dispersion(ctx, step, steps: int) -> dict:
    """Device time of each step on its own (sfn_call_timing event pair)."""
    lib = ctx._lib
    ns = ctypes.c_uint64()
    samples = []
    ctx.check(lib.sfn_call_timing(ctx.ptr, 1))
    try:
        for _ in range(max(steps, 2)):
            step()
            ctx.check(lib.sfn_last_device_ns(ctx.ptr, ctypes.byref(ns)))
            samples.append(ns.value / 1e6)
    finally:
        ctx.check(lib.sfn_call_timing(ctx.ptr, 0))
    res = dispersion(samples)
    res["source"] = "hipEvent pair around each step, device time"
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log-n", type=int, default=20)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--added-bits", type=int, default=1)
    ap.add_argument("--semantics", choices=["as-stated", "reference"], default="as-stated")
    ap.add_argument("--order", choices=["bitrev", "natural"], default="bitrev",
                    help="row order of input and output (natural: sfn_coset_lde_natural, Plonky3's order)")
    ap.add_argument("--cpu-cols", type=int, default=24, help="columns in the CPU baseline sample (0 = skip)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run one 1 GiB device copy (HBM counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--workload", choices=["lde", "bb16", "m31_20", "kb16", "batch20", "uneven", "small"],
                    default="lde")
    ap.add_argument("--e2e-steps", type=int, default=3,
                    help="timed host-pointer (PCIe-inclusive) LDE calls on pinned buffers (0 = skip)")
    ap.add_argument("--cpu-all-cols", type=int, default=128,
                    help="columns in the all-cores CPU baseline sample (0 = skip)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not offered "
              "(the LDE would shard by column, with no exchange between GPUs)", file=sys.stderr)
        return 2

    import numpy as np
    import sfntt
    from sfntt import synth

    log_n, B, a = args.log_n, args.batch, args.added_bits
    n, N = 1 << log_n, 1 << (log_n + a)
    row_major = args.semantics == "as-stated"
    layout = sfntt.LAYOUT_ROW_MAJOR if row_major else sfntt.LAYOUT_SOA
    shift_raw = (31 << 32) % synth.BB_P if row_major else 0

    ctx = sfntt.HipContext(0)
    lib = ctx._lib
    info = ctx.device_info()
    if args.copy_probe:
        # identity transform (log_n = 0) out of place = one plain copy kernel:
        # reads 1 GiB and writes 1 GiB
        cin, cout = sfntt.DeviceBuffer(ctx, 1 << 28), sfntt.DeviceBuffer(ctx, 1 << 28)
        ctx.check(lib.sfn_ntt_device(ctx.ptr, 1, sfntt.VARIANT_AUTO, 0, 0, 1 << 28, sfntt.LAYOUT_SOA,
                                     cin.ptr, cout.ptr, None))
        ctx.check(lib.sfn_synchronize(ctx.ptr, None))
        cin.free()
        cout.free()
    if args.workload != "lde":
        import workloads  # small-field-metal-ntt_amd/workloads.py
        res = workloads.run(args.workload, ctx, args.steps, args.warmup)
        res.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "vs_baseline": None})
        if args.cpu_cols > 0 and args.workload in ("bb16", "m31_20", "kb16"):
            res["cpu_baseline"] = cpu_baseline_ntt(args.workload)
        print(json.dumps(res), flush=True)
        return 0
    # reference generation order: column-major stream, seed 42 + log_n (benches/ntt_benchmark.rs:482-490)
    x = synth.field_data(1, 42 + log_n, n * B)
    x_dev = x.reshape(B, n).T.copy().ravel() if row_major else x
    d_in = sfntt.DeviceBuffer(ctx, n * B)
    d_out = sfntt.DeviceBuffer(ctx, N * B)
    d_in.upload(x_dev)

    lde_fn = lib.sfn_coset_lde_natural_device if args.order == "natural" else lib.sfn_coset_lde_batch_device

    def step():
        ctx.check(lde_fn(ctx.ptr, log_n, B, a, shift_raw, layout, d_in.ptr, d_out.ptr, None))

    for _ in range(args.warmup):
        step()
    ctx.check(lib.sfn_synchronize(ctx.ptr, None))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.check(lib.sfn_synchronize(ctx.ptr, None))
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = N * B * args.steps / elapsed / 1e6

    # dispersion: each step alone between a hipEvent pair on the launch stream
    # (the reference reports median, min, max, stddev, p95 and CV over its
    # timed runs and flags CV above 10 %, benches/ntt_benchmark.rs:75-120)
    disp = step_dispersion(ctx, step, args.steps)

    # per-kernel durations with hipEvent pairs on the launch stream
    lib.sfn_pass_timing(ctx.ptr, 1)
    per = {}
    seq_ns = []
    buf = (ctypes.c_uint64 * 16)()
    cnt = ctypes.c_uint32()
    for _ in range(args.steps):
        step()
        ctx.check(lib.sfn_pass_times(ctx.ptr, buf, 16, ctypes.byref(cnt)))
        tot = 0
        for i in range(cnt.value):
            lab = lib.sfn_pass_label(ctx.ptr, i).decode()
            per.setdefault(lab, []).append(buf[i])
            tot += buf[i]
        seq_ns.append(tot)
    lib.sfn_pass_timing(ctx.ptr, 0)
    kernels = {k: sum(v) / len(v) / 1e3 for k, v in per.items()}  # us
    dom = max(kernels, key=kernels.get)
    dom_bytes = kernel_bytes(dom, log_n, B, a)
    achieved = dom_bytes / (kernels[dom] * 1e-6) / 1e9
    traffic = read_traffic(traffic_key(dom))
    valu_insts = read_valu(traffic_key(dom))
    lde_bytes = 4 * B * (n + N)
    pipe_gbs = lde_bytes / (ms_per_step * 1e-3) / 1e9

    out_first = None
    cb = None
    if args.cpu_cols > 0 and args.order == "bitrev":
        cols = min(args.cpu_cols, B)
        y = d_out.download()
        out_first = (y.reshape(N, B).T[:cols].ravel() if row_major else y[:cols * N])
        cb = cpu_baseline(x[:cols * n], log_n, cols, a, shift_raw, out_first)
    cb_all = None
    if args.cpu_all_cols > 0 and args.order == "bitrev":
        cols = min(args.cpu_all_cols, B)
        cb_all = cpu_baseline_all_cores(x[:cols * n], log_n, cols, a, shift_raw)
    e2e = None
    if args.e2e_steps > 0 and args.order == "bitrev":
        d_in.free()
        d_out.free()
        e2e = end_to_end(ctx, x_dev, log_n, B, a, shift_raw, layout, args.e2e_steps)

    line = {
        "metric": "coset-LDE Melem/s (output elements) + % HBM roofline",
        "value": round(value, 2),
        "unit": "Melem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (reference LCG, seed 42+log_n, BabyBear Montgomery form)",
        "config": {
            "workload": f"babybear_coset_lde_2^{log_n}x{B}_{1 << a}x",
            "log_n": log_n, "batch": B, "added_bits": a,
            "layout": "row_major" if row_major else "soa", "row_order": args.order,
            "shift": "field generator 31" if row_major else "two_adic_generator(log_n+added_bits+1)",
            "device": info["name"] or info["gcn_arch"],
        },
        "roofline": {
            "bound": "hbm", "kernel": dom,
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            # HBM bytes per launch of this kernel: rocprofv3 --pmc counters of the
            # same workload, read from the committed summary (not measured in this run)
            "traffic_source": os.path.relpath(COUNTERS_FILE, ROOT) if traffic is not None else None,
        },
        # the same kernel against the VALU issue peak (the bound it is closest to)
        "roofline_valu": None if valu_insts is None else {
            "bound": "valu", "kernel": dom, "insts_per_launch": valu_insts,
            "achieved": round(valu_insts * 64 / (kernels[dom] * 1e-6) / 1e12, 2),
            "peak": round(VALU_PEAK_WAVE_INSTS * 64 / 1e12, 2), "unit": "Tlane-op/s",
            "frac": round(valu_insts / (kernels[dom] * 1e-6) / VALU_PEAK_WAVE_INSTS, 4),
        },
        "pipeline": {
            "algorithmic_bytes": lde_bytes,
            "achieved_GBs": round(pipe_gbs, 1),
            "frac_hbm": round(pipe_gbs / HBM_PEAK_GBS, 4),
            "device_ms_events": round(sum(seq_ns) / len(seq_ns) / 1e6, 4),
            "kernels_us": {k: round(v, 1) for k, v in kernels.items()},
        },
        "dispersion": disp,
        "cpu_baseline": cb,
        "cpu_baseline_all_cores": cb_all,
        "end_to_end": e2e,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
