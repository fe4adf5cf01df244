"""Serving benchmark: open-loop Poisson arrivals of mixed NTT / coset-LDE jobs.

Job mix (seeded): 60% BabyBear coset LDE (2^14..2^20 rows x 8..256 columns, 2x or
4x blowup), 40% M31 circle NTT (2^12..2^22, one column).

Two systems on the same GPU and the same arrival trace, at two data boundaries:
  plain  - one job per launch sequence, in arrival order, synchronous (the
           reference's calling pattern: one call, wait, next call);
  server - the continuous-batching loop (sfn_server_*): admission by HBM
           footprint, several worker streams, NTT jobs of any shape and LDE
           jobs of one shape fused, large jobs one at a time beside a lane
           for small ones.
Both run in the native load generator serve_load (built beside the library),
so the driver adds microseconds per job, not the Python interpreter's tens.
Data boundaries:
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
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


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


LOADGEN = os.path.join(HERE, "sfntt", "serve_load")


def write_trace(path, jobs, arrival_sets):
    """The load generator's trace file (format in serve_load.cpp)."""
    import struct
    with open(path, "wb") as fh:
        fh.write(struct.pack("<II", len(jobs), len(arrival_sets)))
        for kind, n, b, a in jobs:
            fh.write(struct.pack("<6I", 1 if kind == "lde" else 0, 1 if kind == "lde" else 0, n, b, a, 0))
        for arr in arrival_sets:
            fh.write(np.asarray(arr, dtype="<f8").tobytes())


def run_load(jobs, arrival_sets, system: str, mode: str, inflight: int, group: int):
    """Replays every arrival sequence through the native load generator
    (serve_load.cpp): one process, the shapes warmed once; returns
    [(latencies in s, span in s, mean group)] per sequence."""
    import subprocess
    import tempfile
    with tempfile.NamedTemporaryFile(suffix=".trace", delete=False) as fh:
        path = fh.name
    write_trace(path, jobs, arrival_sets)
    try:
        r = subprocess.run([LOADGEN, path, system, mode, str(inflight), str(group)], capture_output=True, text=True,
                           check=True)
    finally:
        os.unlink(path)
    out = []
    for line in r.stdout.splitlines():
        d = json.loads(line)
        out.append((np.array(d["lat_us"]) / 1e6, d["span_s"], d["mean_group"]))
    return out


def sweep(jobs, mode, inflight, group, rates_arg, gaps):
    """Plain and server over the same arrival traces; rates scale the plain capacity."""
    ((_, span0, _),) = run_load(jobs, [np.zeros(len(jobs))], "plain", mode, inflight, group)
    cap_plain = len(jobs) / span0
    rates = [float(r) for r in rates_arg.split(",")] if rates_arg else \
        [round(cap_plain * f, 2) for f in (0.3, 0.5, 0.7, 0.85, 1.0, 1.2, 1.5, 2.0, 3.0, 5.0)]
    arrs = [np.cumsum(gaps / lam) - gaps[0] / lam for lam in rates]
    plain = run_load(jobs, arrs, "plain", mode, inflight, group)
    server = run_load(jobs, arrs, "server", mode, inflight, group)
    rows = []
    for lam, (lp, sp, _), (ls, ss, g) in zip(rates, plain, server):
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
    ap.add_argument("--inflight", type=int, default=6, help="server worker streams")
    ap.add_argument("--group", type=int, default=128, help="most jobs the server fuses into one launch set")
    ap.add_argument("--modes", default="device,host")
    ap.add_argument("--write-trace", default="",
                    help="only write the trace (every job arriving at once, closed loop) to this file, e.g. to "
                         "replay it under a profiler: sfntt/serve_load FILE plain|server device|host")
    args = ap.parse_args()
    modes = [m for m in args.modes.split(",") if m]
    jobs = draw_jobs(0x5EED0005, args.jobs, args.scale)
    if args.write_trace:
        write_trace(args.write_trace, jobs, [np.zeros(len(jobs))])
        return
    rng = np.random.default_rng(0x5EED0005 + 1)
    gaps = rng.exponential(1.0, len(jobs))
    targets = [float(t) for t in args.p99_ms.split(",") if t]
    rows_by_mode, caps = {}, {}
    for mode in modes:
        rows_by_mode[mode], caps[mode] = sweep(jobs, mode, args.inflight, args.group, args.rates, gaps)
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
