"""Serving benchmark: open-loop Poisson arrivals of mixed NTT / coset-LDE jobs.

Job mix (seeded): 60% BabyBear coset LDE (2^14..2^20 rows x 8..256 columns, 2x or
4x blowup), 40% M31 circle NTT (2^12..2^22, one column).  Columns stay resident
in HBM: inputs are device buffers, every job writes its own device output.

Two systems on the same GPU and the same arrival trace, at two data boundaries:
  plain  - one job per launch sequence, in arrival order, synchronous (the
           reference's calling pattern: one call, wait, next call);
  server - the continuous-batching loop (sfn_server_*): FIFO admission by HBM
           footprint, several worker streams, same-shape jobs fused.
  device - columns resident in HBM (inputs are device buffers, every job
           writes its own device output);
  host   - pinned host buffers: every job uploads its input and downloads its
           result (plain: sfn_ntt / sfn_coset_lde_batch; server: requests with
           device_ptrs = 0).  All host-mode jobs read one shared input buffer
           and write one shared output buffer: the timing does not depend on
           the destination, and results are checked by the tests instead.
For each offered rate the p50/p99 latency (arrival -> result complete) and
the completed jobs/s are reported; the headline is the highest rate each
system sustains under a p99 target, and their ratios.

    python small-field-metal-ntt_amd/serve_bench.py [--jobs 200] [--modes device,host] [--p99-ms 100]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import sfntt  # noqa: E402
from sfntt import _lib as L  # noqa: E402
from sfntt import synth  # noqa: E402


def draw_jobs(seed: int, count: int, scale: int = 0):
    """(kind, log_n, batch, added_bits) per job; `scale` shrinks sizes for quick runs."""
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(count):
        if rng.random() < 0.6:
            jobs.append(("lde", int(rng.integers(14, 21)) - scale, int(rng.integers(8, 257)),
                         int(rng.integers(1, 3))))
        else:
            jobs.append(("ntt", int(rng.integers(12, 23)) - scale, 1, 0))
    return jobs


class Pool:
    """Device buffers (one shared read-only input, one output per job) and, for
    host mode, one shared pinned input and one shared pinned output."""

    def __init__(self, ctx, jobs, modes):
        self.ctx = ctx
        max_in = max(b << n for _, n, b, _ in jobs)
        max_out = max(b << (n + a) for _, n, b, a in jobs)
        x = synth.field_data(1, 0x5EED0005, max_in)
        # M31 inputs must be < 2^31 - 1: BabyBear values are < p_BB < p_M31, valid for both
        self.inp = sfntt.DeviceBuffer(ctx, max_in)
        self.inp.upload(x)
        self.outs = [sfntt.DeviceBuffer(ctx, b << (n + a)) for _, n, b, a in jobs] if "device" in modes else []
        self.hin = self.hout = None
        if "host" in modes:
            hin, hout = ctypes.c_void_p(), ctypes.c_void_p()
            ctx.check(ctx._lib.sfn_host_alloc(ctx.ptr, max_in * 4, ctypes.byref(hin)))
            ctx.check(ctx._lib.sfn_host_alloc(ctx.ptr, max_out * 4, ctypes.byref(hout)))
            ctypes.memmove(hin, x.ctypes.data, max_in * 4)
            self.hin, self.hout = hin, hout

    def io(self, mode, k):
        """(input, output) pointers of job k in this mode."""
        if mode == "device":
            return self.inp.ptr, self.outs[k].ptr
        return self.hin, self.hout

    def close(self):
        for h in (self.hin, self.hout):
            if h:
                self.ctx._lib.sfn_host_free(self.ctx.ptr, h)
        self.hin = self.hout = None


def _u32(p):
    return ctypes.cast(p, L.u32p)


def run_plain(ctx, jobs, pool, arrivals, mode="device"):
    lib = ctx._lib
    lat = []
    t0 = time.perf_counter()
    for k, ((kind, n, b, a), at) in enumerate(zip(jobs, arrivals)):
        now = time.perf_counter() - t0
        if now < at:
            time.sleep(at - now)
        src, dst = pool.io(mode, k)
        if mode == "device":
            if kind == "lde":
                st = lib.sfn_coset_lde_batch_device(ctx.ptr, n, b, a, 0, L.LAYOUT_SOA, src, dst, None)
            else:
                st = lib.sfn_ntt_device(ctx.ptr, L.FIELD_M31, L.VARIANT_AUTO, L.FORWARD, n, 1, L.LAYOUT_SOA,
                                        src, dst, None)
            ctx.check(st)
            ctx.check(lib.sfn_synchronize(ctx.ptr, None))
        else:  # the host-pointer calls are synchronous
            if kind == "lde":
                st = lib.sfn_coset_lde_batch(ctx.ptr, n, b, a, 0, L.LAYOUT_SOA, _u32(src), _u32(dst), None)
            else:
                st = lib.sfn_ntt(ctx.ptr, L.FIELD_M31, L.VARIANT_AUTO, L.FORWARD, n, 1, L.LAYOUT_SOA,
                                 _u32(src), _u32(dst), None)
            ctx.check(st)
        lat.append(time.perf_counter() - t0 - at)
    span = time.perf_counter() - t0
    return np.array(lat), span


def run_server(ctx, jobs, pool, arrivals, inflight=4, group=64, mode="device"):
    lib = ctx._lib
    cfg = L.ServerConfig(0, inflight, 0, group, 0)
    srv = ctypes.c_void_p()
    ctx.check(lib.sfn_server_create(ctx.ptr, ctypes.byref(cfg), ctypes.byref(srv)))
    dev = 1 if mode == "device" else 0
    tickets = []
    base = lib.sfn_now()
    t0 = time.perf_counter()
    for k, ((kind, n, b, a), at) in enumerate(zip(jobs, arrivals)):
        now = time.perf_counter() - t0
        if now < at:
            time.sleep(at - now)
        src, dst = pool.io(mode, k)
        if kind == "lde":
            r = L.Request(L.JOB_LDE, L.FIELD_BABYBEAR, 0, n, b, a, 0, dev, src, dst)
        else:
            r = L.Request(L.JOB_NTT, L.FIELD_M31, L.FORWARD, n, 1, 0, 0, dev, src, dst)
        t = ctypes.c_uint64()
        ctx.check(lib.sfn_server_submit(srv, ctypes.byref(r), ctypes.byref(t)))
        tickets.append((t.value, at))
    lat, groups = [], []
    end_max = 0
    for t, at in tickets:
        tm = L.JobTiming()
        ctx.check(lib.sfn_server_wait(srv, t, ctypes.byref(tm)))
        # latency from the scheduled arrival (the submit happens at it)
        lat.append((tm.end_ns - base) / 1e9 - at)
        groups.append(tm.group_size)
        end_max = max(end_max, tm.end_ns)
    lib.sfn_server_destroy(srv)
    return np.array(lat), (end_max - base) / 1e9, float(np.mean(groups))


def sweep(ctx, jobs, pool, mode, inflight, rates_arg, gaps):
    """Plain and server over the same arrival traces; rates scale the plain capacity."""
    lat0, span0 = run_plain(ctx, jobs, pool, np.zeros(len(jobs)), mode)
    cap_plain = len(jobs) / span0
    rates = [float(r) for r in rates_arg.split(",")] if rates_arg else \
        [round(cap_plain * f, 2) for f in (0.3, 0.5, 0.7, 0.85, 1.0, 1.2, 1.5, 2.0, 3.0, 5.0)]
    rows = []
    for lam in rates:
        arr = np.cumsum(gaps / lam) - gaps[0] / lam
        lp, sp = run_plain(ctx, jobs, pool, arr, mode)
        ls, ss, g = run_server(ctx, jobs, pool, arr, inflight, mode=mode)
        row = {"mode": mode, "offered_jobs_s": lam,
               "plain": {"jobs_s": round(len(jobs) / sp, 2), "p50_ms": round(np.percentile(lp, 50) * 1e3, 3),
                         "p99_ms": round(np.percentile(lp, 99) * 1e3, 3)},
               "server": {"jobs_s": round(len(jobs) / ss, 2), "p50_ms": round(np.percentile(ls, 50) * 1e3, 3),
                          "p99_ms": round(np.percentile(ls, 99) * 1e3, 3), "mean_group": round(g, 2)}}
        print(json.dumps(row), flush=True)
        rows.append(row)
    return rows, cap_plain


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=200)
    ap.add_argument("--rates", default="")
    ap.add_argument("--p99-ms", default="20,50,100,200", help="p99 targets (ms), comma separated")
    ap.add_argument("--scale", type=int, default=0, help="shrink every log_n by this much (quick runs)")
    ap.add_argument("--inflight", type=int, default=4)
    ap.add_argument("--modes", default="device,host")
    args = ap.parse_args()
    modes = [m for m in args.modes.split(",") if m]
    ctx = sfntt.HipContext(0)
    jobs = draw_jobs(0x5EED0005, args.jobs, args.scale)
    pool = Pool(ctx, jobs, modes)
    # warm every shape's tables once, outside the timed traces
    for k, (kind, n, b, a) in enumerate(jobs):
        src, dst = pool.inp.ptr, (pool.outs[k].ptr if pool.outs else None)
        if dst is None:
            continue
        if kind == "lde":
            ctx.check(ctx._lib.sfn_coset_lde_batch_device(ctx.ptr, n, b, a, 0, 0, src, dst, None))
        else:
            ctx.check(ctx._lib.sfn_ntt_device(ctx.ptr, 0, 4, 0, n, 1, 0, src, dst, None))
    ctx.check(ctx._lib.sfn_synchronize(ctx.ptr, None))
    rng = np.random.default_rng(0x5EED0005 + 1)
    gaps = rng.exponential(1.0, len(jobs))
    targets = [float(t) for t in args.p99_ms.split(",") if t]
    rows_by_mode, caps = {}, {}
    for mode in modes:
        rows_by_mode[mode], caps[mode] = sweep(ctx, jobs, pool, mode, args.inflight, args.rates, gaps)
    pool.close()
    for target in targets:
        best = {}
        for mode, rows in rows_by_mode.items():
            for sys_ in ("plain", "server"):
                ok = [r for r in rows if r[sys_]["p99_ms"] <= target]
                best[f"{sys_}_{mode}"] = max((r[sys_]["jobs_s"] for r in ok), default=0.0)

        def ratio(a, b):
            return round(best[a] / best[b], 3) if best.get(a) and best.get(b) else None

        print(json.dumps({"metric": "serving jobs/s at p99 <= target", "p99_target_ms": target,
                          "jobs_per_trace": len(jobs), "scale": args.scale,
                          "best_jobs_s": best,
                          "ratio_server_vs_plain_device": ratio("server_device", "plain_device"),
                          "ratio_server_vs_plain_host": ratio("server_host", "plain_host"),
                          "ratio_server_device_vs_plain_host": ratio("server_device", "plain_host"),
                          "plain_capacity_jobs_s": {m: round(c, 2) for m, c in caps.items()}}), flush=True)


if __name__ == "__main__":
    main()
