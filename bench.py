#!/usr/bin/env python3
"""Benchmark of the MI355X shuffle hash split (and, with --workload, the other
hot-path operators).

Default workload: shuffle hash split of 64M rows in 4096-row batches
(int64 key U[0,2^40), int32, float64 N(0,1) with 2% nulls, utf8 of length
U[8,32] over [a-z]) into 200 partitions by Spark Murmur3 (seed 42) pmod.  A
step splits all 16384 batches in one batched launch sequence; inputs are
resident in HBM before timing starts.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=64 * 1024 * 1024)
    ap.add_argument("--plain-batches", type=int, default=2048,
                    help="batches timed on the plain per-batch path")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target CPU time of the baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-rows", type=int, default=0, help="split tile rows (0 = default)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="also run a device copy of this many MiB (PMC calibration)")
    return ap.parse_args()


def load_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), committed under profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def cpu_baseline_split(target_s: float):
    """The oracle's per-batch reference split loop on one host core."""
    import numpy as np

    import oracle
    from gazelle_plugin_amd import datagen

    rows = 256 * 4096
    batches = datagen.split_host(rows)
    cb = oracle._CBatches(batches)
    sp = oracle.CpuSplitter(datagen.SPLIT_TYPES, 200, [0])
    t0 = time.perf_counter()
    for i in range(len(batches)):
        sp.split(cb, i)
    t1 = time.perf_counter()
    sp.close()
    per_row = (t1 - t0) / rows
    # extend the sample to about target_s of work by repeating the pass
    reps = max(1, int(target_s / max(1e-9, t1 - t0)))
    total_rows = rows
    total_t = t1 - t0
    for _ in range(reps - 1):
        sp = oracle.CpuSplitter(datagen.SPLIT_TYPES, 200, [0])
        t0 = time.perf_counter()
        for i in range(len(batches)):
            sp.split(cb, i)
        total_t += time.perf_counter() - t0
        total_rows += rows
        sp.close()
    return {"value": total_rows / total_t, "unit": "rows/s", "cores": 1, "kind": "port",
            "sample": f"{total_rows} rows ({reps} passes over 256 batches x 4096 rows) of the "
                      f"same schema, oracle per-batch split loop (splitter.cc DoSplit restated), "
                      f"{total_t:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is "
                                   "described in DESIGN.md but not launched", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from gazelle_plugin_amd import datagen
    from gazelle_plugin_amd.context import Context
    from gazelle_plugin_amd.splitter import Splitter

    torch.cuda.init()
    stream = torch.cuda.current_stream()
    ctx = Context(0, stream=stream.cuda_stream)
    tab = datagen.split_device(args.rows)
    torch.cuda.synchronize()
    N = tab.num_rows
    in_bytes = tab.nbytes
    P = 200
    out_bytes = in_bytes + 4 * P  # same columns out, offsets gain one entry per partition
    algo_bytes = in_bytes + out_bytes

    from gazelle_plugin_amd.splitter import make_batch_array

    arr = make_batch_array(tab.batches)

    from gazelle_plugin_amd.splitter import SplitOptions

    opts = SplitOptions(tile_rows=args.tile_rows)

    def step():
        sp = Splitter.Make(ctx, "hash", datagen.SPLIT_TYPES, P, [0], opts)
        sp.SplitBatches(tab.batches, arr)
        sp.close()  # buffers return to the pool in stream order (no host sync)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.set_profiling(True)
    ctx.reset_kernel_times()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = gpu_ms / args.steps
    scat_ms, scat_n = ctx.kernel_time("split_scatter")
    cnt_ms, cnt_n = ctx.kernel_time("split_count")
    scan_ms, scan_n = ctx.kernel_time("split_scan")
    ctx.set_profiling(False)
    scat_avg = scat_ms / max(1, scat_n)
    rows_per_s = N / (ms_per_step / 1e3)

    # plain per-batch path on a prefix of the batches
    nb = min(args.plain_batches, len(tab.batches))
    sp = Splitter.Make(ctx, "hash", datagen.SPLIT_TYPES, P, [0])
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for b in tab.batches[:nb]:
        sp.Split(b)
    e1.record(stream)
    torch.cuda.synchronize()
    plain_ms = e0.elapsed_time(e1)
    sp.close()
    plain_rows_s = nb * datagen.BATCH_ROWS / (plain_ms / 1e3)

    traffic = load_traffic("split_scatter_traffic.json")
    roof = {"bound": "hbm", "achieved": algo_bytes / (scat_avg / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
            "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = traffic.get("bytes_per_launch") if traffic else None
    roof["kernel"] = "split_scatter"
    roof["algorithmic_bytes"] = algo_bytes
    roof["kernel_ms"] = scat_avg

    if args.calibrate_copy:
        # a plain copy of known size: the PMC FETCH_SIZE / WRITE_SIZE it reads
        # scale the split kernel's counters to bytes
        src = torch.empty(args.calibrate_copy << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
    cpu = None if args.no_cpu else cpu_baseline_split(args.cpu_seconds)
    out = {
        "metric": "rows/s + HBM GB/s (% peak), batched vs plain GPU path; p50/p99 fragment latency",
        "value": rows_per_s,
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded, generated in HBM)",
        "config": {"workload": "shuffle hash split: 64M rows in 4096-row batches, int64 key "
                               "U[0,2^40) + int32 + float64 (2% nulls) + utf8 len U[8,32], "
                               "200 partitions, Spark Murmur3 seed 42 pmod",
                   "rows": N, "batches": len(tab.batches), "partitions": P,
                   "path": "batched (all batches in one launch sequence)"},
        "step_GBps": algo_bytes / (ms_per_step / 1e3) / 1e9,
        "kernels_ms_avg": {"split_count": cnt_ms / max(1, cnt_n), "split_scan": scan_ms / max(1, scan_n),
                           "split_scatter": scat_avg},
        "plain": {"rows_per_s": plain_rows_s, "batches": nb, "ms": plain_ms},
        "batched_over_plain": rows_per_s / plain_rows_s,
        "wall_s": t1 - t0,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
