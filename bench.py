"""Benchmark of the asynchronous execution runtime on one MI355X.

Workload: Collatz at scale. 2^26 uint64 starts drawn uniformly from [1, 2^40)
by Philox4x32-10 (seed 0x5EED), an arena of 2^24 promises (2^18 links of 64),
uint32 step counts out. A step is one complete AsyncProgram run over all
starts (init + exec launches until complete); the starts stay resident in HBM.

Prints one JSON line: promises resolved per second of AsyncProgram, with the
EventProgram and one-thread-per-start comparators, the roofline of the
dominant kernel and the CPU baseline (the oracle's host executors on a
bounded sample of the same workload).

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--no-cpu] [--no-event]
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "promises resolved/sec (Async vs Event, 1 GPU); p50/p99 job latency"
SEED = 0x5EED
BITS = 40
COUNT = 1 << 26
ARENA_PROMISES = 1 << 24
LINK_PROMISES = 64
HBM_PEAK_GBPS = 8000.0
PROMISE_BYTES = 16  # packed Collatz promise {u64 value, u32 step, u32 index}


def algorithmic_bytes(promises: int, starts: int) -> int:
    """Enqueue write + dequeue read of every promise, plus start reads and step writes."""
    return 2 * PROMISE_BYTES * promises + starts * (8 + 4)


def cpu_baseline(threads_all: int) -> dict:
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O  # the oracle, used only as the CPU baseline leg

    sample = 1 << 22
    starts = O.collatz_make_starts(SEED, BITS, COUNT)[:sample]
    _, st_omp = O.collatz_program(starts, threads=0)
    omp_rate = st_omp.promises / (st_omp.elapsed_ms * 1e-3)
    seq_sample = 1 << 20
    _, st_seq = O.collatz_program(starts[:seq_sample], threads=1)
    seq_rate = st_seq.promises / (st_seq.elapsed_ms * 1e-3)
    return {
        "value": omp_rate,
        "unit": "promises/s",
        "cores": int(st_omp.threads),
        "kind": "port",
        "sample": f"first {sample} of the 2^26 starts, Collatz spec on the oracle's OpenMP host executor "
                  f"({st_omp.promises} promises, {st_omp.elapsed_ms / 1e3:.1f} s)",
        "sequential": {
            "value": seq_rate,
            "cores": 1,
            "sample": f"first {seq_sample} starts on the sequential host executor "
                      f"({st_seq.promises} promises, {st_seq.elapsed_ms / 1e3:.1f} s)",
        },
        "cpu_model": platform.processor() or platform.machine(),
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--groups", type=int, default=0, help="AsyncProgram groups (0 = fill the device)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-event", action="store_true", help="skip the EventProgram comparator")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this runtime is measured on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "described in DESIGN.md, not built)"}))
        return 2

    import harmonize_amd as hz

    facts = hz.device_facts()
    starts = hz.collatz_make_starts(SEED, BITS, COUNT)
    steps_out = hz.DeviceArray(COUNT, np.uint32)
    arena_links = ARENA_PROMISES // LINK_PROMISES
    cfg = dict(arena=arena_links, groups=args.groups)

    for _ in range(args.warmup):
        hz.collatz_run("async", steps_out, starts, COUNT, **cfg)
    hz.check(hz.lib().hz_device_synchronize(), "sync")

    runs = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        runs.append(hz.collatz_run("async", steps_out, starts, COUNT, **cfg))
    hz.check(hz.lib().hz_device_synchronize(), "sync")
    wall = time.perf_counter() - t0

    promises = runs[-1]["promises"]
    total_promises = sum(r["promises"] for r in runs)
    value = total_promises / wall
    kernel_ms = float(np.mean([r["elapsed_ms"] for r in runs]))
    launches = int(np.mean([r["launches"] for r in runs]))
    per_launch_ms = kernel_ms / max(1, launches)
    alg = algorithmic_bytes(promises, COUNT) / max(1, launches)
    achieved = alg / (per_launch_ms * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "promises/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": "collatz_at_scale: 2^26 u64 starts uniform in [1,2^40) (Philox4x32-10 seed 0x5EED), "
                        "arena 2^24 promises, u32 steps",
            "program": "AsyncProgram",
            "groups": runs[-1]["groups"],
            "arena_links": arena_links,
            "promises_per_step": promises,
            "launches_per_step": launches,
            "spills": runs[-1]["spills"],
            "fills": runs[-1]["fills"],
            "device": facts["arch"],
            "compute_units": facts["compute_units"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "kernel": "_hz_exec_kernel<AsyncProgram<CollatzSpec>>",
            "kernel_ms": per_launch_ms,
            "bytes_rule": "2 x 16 B per promise (enqueue + dequeue of the packed payload) + 12 B per start",
        },
    }

    if not args.no_event:
        ev = hz.collatz_run("event", steps_out, starts, COUNT, max_launches=1 << 20)
        out["event"] = {
            "value": ev["promises"] / (ev["elapsed_ms"] * 1e-3),
            "unit": "promises/s",
            "ms": ev["elapsed_ms"],
            "launches": ev["launches"],
            "groups": ev["groups"],
        }
        out["async_vs_event"] = value / out["event"]["value"]

    pl = hz.collatz_run("plain", steps_out, starts, COUNT)
    out["plain_kernel"] = {
        "ms": pl["elapsed_ms"],
        "steps_per_s": promises / (pl["elapsed_ms"] * 1e-3),
        "note": "one thread per start, no runtime (steps computed per second, same work)",
    }

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(os.cpu_count() or 1)

    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
