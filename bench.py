"""Benchmark of the asynchronous execution runtime on one MI355X.

Workload: Collatz at scale. 2^26 uint64 starts drawn uniformly from [1, 2^40)
by Philox4x32-10 (seed 0x5EED), an arena of 2^24 promises (2^18 links of 64),
uint32 step counts out. A step is one complete AsyncProgram run over all
starts (init + exec launches until complete); the starts stay resident in HBM.

Prints one JSON line: promises resolved per second of AsyncProgram, with the
EventProgram and one-thread-per-start comparators, the roofline of the
dominant kernel and the CPU baseline (the oracle's host executors on a
bounded sample of the same workload). Further sections: the photon walk and
the batched BFS (Async vs Event), the futures tree, the remapping-barrier demo
and the serving load sweep (p50/p99 latency at 20-100% of peak).

Consistency checks made here (a failure raises): the Collatz step counts of
AsyncProgram, EventProgram and the one-thread-per-start kernel are identical;
photon event totals and BFS levels agree between the two programs; the futures
total and the barrier demo's count match their closed forms. The GPU tests
check every workload, serving included, against the oracle.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--no-cpu] [--no-event]
                    [--no-photon] [--no-bfs] [--no-futures] [--no-barrier] [--no-serve]
    python bench.py --leg <name>   # one leg once, for counter collection
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
QUEUES = 16  # link queues of the async program (the tuner's choice; 8 is 1% slower)


def algorithmic_bytes(promises: int, starts: int) -> int:
    """Enqueue write + dequeue read of every promise, plus start reads and step writes."""
    return 2 * PROMISE_BYTES * promises + starts * (8 + 4)


# HBM traffic of one AsyncProgram exec launch on this workload, from two
# `rocprofv3 --pmc` passes over `bench.py --leg collatz_async` (FETCH_SIZE,
# then WRITE_SIZE, KiB per dispatch), each scaled by what the same counter
# reads for a 1 GiB device-to-device copy (`bench.py --leg copy`): FETCH_SIZE
# reads 512 MiB for the copy's 1 GiB of reads (factor 2.0) and WRITE_SIZE
# reads 1 GiB for its 1 GiB of writes (factor 1.0).
TRAFFIC_KIB = {"fetch": 1011086.5, "write": 2198988.1}
TRAFFIC_FACTOR = {"fetch": 2.0, "write": 1.0}


def measured_traffic_bytes() -> int:
    return int(sum(TRAFFIC_KIB[k] * TRAFFIC_FACTOR[k] * 1024 for k in TRAFFIC_KIB))


# Instruction counts of one AsyncProgram exec launch on this workload, from a
# `rocprofv3 --pmc SQ_INSTS_SALU SQ_INSTS_VALU SQ_INSTS_LDS SQ_INSTS_BRANCH
# SQ_WAVES SQ_BUSY_CYCLES GRBM_GUI_ACTIVE GRBM_COUNT` pass over
# `bench.py --leg collatz_async` (totals over all wavefronts of the dispatch).
ISSUE_COUNTS = {"salu": 2.4486e10, "valu": 1.8018e10, "lds": 4.1635e9, "branch": 8.7764e9}
CLOCK_GHZ = 2.4  # MI355X peak engine clock
LDS_BYTES_PER_CLOCK_PER_CU = 128


_T0 = time.perf_counter()


def progress(what: str) -> None:
    """A line on stderr after each part of the run (stdout carries only the
    result line)."""
    print(f"bench: {what} done at {time.perf_counter() - _T0:.1f} s", file=sys.stderr, flush=True)


def cpu_model() -> str:
    """The host CPU's model name (from /proc/cpuinfo, as lscpu reports it)."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(repeats: int = 5) -> dict:
    """The oracle's host executors on a bounded sample of the headline workload:
    one warm-up, then the median of `repeats` timed runs each (steady_clock
    inside the executor)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O  # the oracle, used only as the CPU baseline leg

    def median_rate(starts, threads):
        O.collatz_program(starts, threads=threads)
        runs = [O.collatz_program(starts, threads=threads)[1] for _ in range(repeats)]
        rates = sorted(r.promises / (r.elapsed_ms * 1e-3) for r in runs)
        return rates[len(rates) // 2], runs[0]

    sample = 1 << 22
    starts = O.collatz_make_starts(SEED, BITS, COUNT)[:sample]
    omp_rate, st_omp = median_rate(starts, 0)
    seq_sample = 1 << 19
    seq_rate, st_seq = median_rate(starts[:seq_sample], 1)
    return {
        "value": omp_rate,
        "unit": "promises/s",
        "cores": int(st_omp.threads),
        "kind": "port",
        "sample": f"first {sample} of the 2^26 starts, Collatz spec on the oracle's OpenMP host executor "
                  f"({st_omp.promises} promises, {st_omp.elapsed_ms / 1e3:.2f} s per run); "
                  f"median of {repeats} runs after one warm-up",
        "sequential": {
            "value": seq_rate,
            "cores": 1,
            "sample": f"first {seq_sample} starts on the sequential host executor "
                      f"({st_seq.promises} promises, {st_seq.elapsed_ms / 1e3:.2f} s per run); "
                      f"median of {repeats} runs after one warm-up",
        },
        "cpu_model": cpu_model(),
        "host_cpus": os.cpu_count(),
        "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
        "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
    }


def photon_comparison(hz, repeats: int = 3) -> dict:
    """Branching photon walk, 1e7 photons: AsyncProgram vs EventProgram, same spec,
    each the best of `repeats` runs at its best measured launch shape."""
    p = hz.photon_params()
    tally = hz.DeviceArray(p.replicas * (p.bins + hz.PHOTON_EXTRA_SLOTS), np.float64)
    events = hz.DeviceArray(p.photons, np.uint32)

    def best(kind, **cfg):
        runs = [hz.photon_run(kind, p, tally, events, **cfg)[2] for _ in range(repeats)]
        return min(runs, key=lambda r: r["elapsed_ms"])

    a = best("async")
    e = best("event", groups=2048, cycles=8, max_launches=1 << 20)
    if a["promises"] != e["promises"]:
        raise RuntimeError("photon: async and event evaluated different numbers of events")
    rate = lambda r: r["promises"] / (r["elapsed_ms"] * 1e-3)
    return {
        "workload": "branching photon walk: 1e7 photons, slab 20, sigma_t 1, albedo 0.9, split > 2, "
                    "roulette < 0.1, Philox4x32-10, 1000 FP64 depth bins",
        "events": a["promises"],
        "async": {"value": rate(a), "unit": "events/s", "ms": a["elapsed_ms"], "groups": a["groups"]},
        "event": {"value": rate(e), "unit": "events/s", "ms": e["elapsed_ms"], "groups": e["groups"],
                  "launches": e["launches"]},
        "async_vs_event": rate(a) / rate(e),
    }


FUTURES_LEAVES = 1 << 20
FUTURES_READERS = 1 << 14


def futures_comparison(hz, repeats: int = 3) -> dict:
    """Value-returning operations: a tree of 2^20 unit futures summing the Collatz
    step counts of 1..2^20 (each leaf a tail-called chain of Steps promises),
    resolving a RemappingBarrier future that 2^14 readers await and forward.
    EventProgram runs the tree only (it cannot hold promises on a barrier)."""
    starts = hz.DeviceArray.from_numpy(np.arange(1, FUTURES_LEAVES + 1, dtype=np.uint64))
    expected_total = 139_348_407 - FUTURES_LEAVES  # the reference checker's sum of (steps + 1) over 1..2^20

    def best(kind, readers):
        runs = []
        for _ in range(repeats):
            total, echo, forwarded, st = hz.futures_run(kind, starts, readers)
            if total != expected_total or (readers and (int(echo[-1]) != total + readers - 1 or
                                                        int(forwarded.max()) != total)):
                raise RuntimeError(f"futures ({kind}): wrong result {total}")
            runs.append(st)
        return min(runs, key=lambda r: r["elapsed_ms"])

    a = best("async", FUTURES_READERS)
    a_tree = best("async", 0)
    e = best("event", 0)
    rate = lambda r: r["promises"] / (r["elapsed_ms"] * 1e-3)
    return {
        "workload": "futures tree: 2^20 leaves (Collatz steps of 1..2^20 through tail-called value-returning "
                    "promises), unit-future joins, a RemappingBarrier total awaited and forwarded by 2^14 readers",
        "async": {"value": rate(a), "unit": "promises/s", "ms": a["elapsed_ms"], "promises": a["promises"],
                  "readers": FUTURES_READERS},
        "async_tree_only": {"value": rate(a_tree), "unit": "promises/s", "ms": a_tree["elapsed_ms"],
                            "promises": a_tree["promises"]},
        "event_tree_only": {"value": rate(e), "unit": "promises/s", "ms": e["elapsed_ms"],
                            "promises": e["promises"], "launches": e["launches"]},
        "async_vs_event_tree": rate(a_tree) / rate(e),
    }


BARRIER_DEP, BARRIER_NON_DEP = 1 << 21, 1 << 21


def barrier_demo(hz, repeats: int = 3) -> dict:
    """The reference's remapping-barrier demo (examples/cpp/barrier_demo/demo.cpp)
    at scale: 2^21 dependent steps, each awaiting barrier[id % barriers] before
    it runs again, and 2^21 independent steps. Checked as the demo checks it
    (counter = 2 * dependents + independents, demo.cpp:214), plus every held
    step run exactly once and none before its barrier's release. Measured with
    1,024 barriers (2,048 awaits each) and with 64 (32,768 awaits each), where
    each barrier's awaits share one ticket counter and its release queues 512
    held links."""

    def best(barriers, reps):
        runs = []
        for _ in range(reps):
            counter, hits, violations, st = hz.barrier_run(BARRIER_DEP, BARRIER_NON_DEP, barriers, arena=1 << 17,
                                                           max_launches=1000)
            if counter != 2 * BARRIER_DEP + BARRIER_NON_DEP or violations != 0 or not bool((hits == 1).all()):
                raise RuntimeError(f"barrier demo ({barriers} barriers): counter {counter}, violations {violations}")
            runs.append(st)
        return min(runs, key=lambda r: r["elapsed_ms"])

    out = {"workload": "remapping-barrier demo: 2^21 dependent steps awaiting barrier[id % barriers], "
                       "2^21 independent steps (AsyncProgram; EventProgram cannot hold promises on a barrier)"}
    for barriers, reps in ((1024, repeats), (64, 1)):
        r = best(barriers, reps)
        out[f"barriers_{barriers}"] = {
            "ms": r["elapsed_ms"],
            "held_promises_per_s": r["promises"] / (r["elapsed_ms"] * 1e-3),
            "steps_per_s": (2 * BARRIER_DEP + BARRIER_NON_DEP) / (r["elapsed_ms"] * 1e-3),
            "held_links": r["links_run"],
        }
    return out


BFS_SEED = 0x5EED
BFS_GRAPHS = 64
# Autotuner results per program, device and workload class (hz_tune_cache_*).
TUNE_CACHE = ROOT / "harmonize_amd" / "tune_cache.txt"


def bfs_comparison(hz, repeats: int = 3) -> dict:
    """64 R-MAT graphs (2^14..2^22 vertices, 16 edges per vertex) searched by
    one program: AsyncProgram through its captured hipGraph, at the
    configuration the autotuner picks for this batch, vs EventProgram, each the
    best of `repeats` runs. Levels must agree between the two."""
    sizes = hz.bfs_batch_sizes(BFS_SEED, BFS_GRAPHS, 14.0, 22.0)
    batch = hz.BfsBatch.rmat(BFS_SEED, sizes)
    try:
        def best(kind, **cfg):
            runs = []
            for _ in range(repeats):
                runs.append(batch.run(kind, max_launches=1 << 16, **cfg))
            return min(runs, key=lambda r: r["elapsed_ms"]), batch.levels()

        # The tuner's choice for this batch class on this device comes from
        # the shipped cache when present (no trials); otherwise it tunes and
        # the cache records the result.
        hz.tune_cache_set_path(str(TUNE_CACHE))
        try:
            tuned = hz.tune_bfs(batch, rounds=2, reps=2)
        finally:
            hz.tune_cache_set_path(None)
        cfg = {k: v for k, v in tuned["best"].items() if v and k != "max_launches"}
        cfg["flags"] = hz.RUN_GRAPH
        untuned, _ = best("async", flags=hz.RUN_GRAPH)
        a, lev_a = best("async", **cfg)
        e, lev_e = best("event")
        if not np.array_equal(lev_a, lev_e):
            raise RuntimeError("bfs: async and event levels differ")
        reached = lev_a != hz.BFS_UNREACHED
        traversed = traversed_edges(hz, batch, reached)
        rate = lambda r: r["promises"] / (r["elapsed_ms"] * 1e-3)
        return {
            "workload": f"{BFS_GRAPHS} R-MAT graphs (a,b,c = 0.57,0.19,0.19), vertices log-uniform in [2^14, 2^22] "
                        f"(seed 0x5EED), 16 edges per vertex, int32 CSR, BFS from vertex 0 of each",
            "vertices": int(batch.total_vertices),
            "edges": int(batch.total_edges),
            "reached": int(reached.sum()),
            "traversed_edges": traversed,
            "build_ms": batch.build_ms,
            "async": {"visits": a["promises"], "ms": a["elapsed_ms"], "launches": a["launches"],
                      "groups": a["groups"], "promises_per_s": rate(a), "teps": traversed / (a["elapsed_ms"] * 1e-3),
                      "launch": "hipGraph", "tuned_config": cfg, "untuned_ms": untuned["elapsed_ms"],
                      "tuning_trials": tuned["trials"], "tuning_cached": tuned["cached"]},
            "event": {"visits": e["promises"], "ms": e["elapsed_ms"], "launches": e["launches"],
                      "groups": e["groups"], "promises_per_s": rate(e), "teps": traversed / (e["elapsed_ms"] * 1e-3)},
            "async_vs_event_time": e["elapsed_ms"] / a["elapsed_ms"],
        }
    finally:
        batch.close()


def traversed_edges(hz, batch, reached) -> int:
    """Sum of out-degrees of reached vertices (edges a BFS must scan)."""
    row_ptr, _ = batch.csr_rows()
    total = 0
    for g in range(batch.graphs):
        v0, n = int(batch.vertex_base[g]), int(batch.vertices[g])
        r0 = int(batch.row_base[g])
        deg = np.diff(row_ptr[r0:r0 + n + 1].astype(np.int64))
        total += int(deg[reached[v0:v0 + n]].sum())
    return total


SERVE_JOBS = 400


def serving_mix(n: int, seed: int) -> list:
    """Jobs drawn as the serving workload defines them: Collatz, photon walk or
    BFS with probability 1/3 each, size log-uniform in [2^10, 2^22] work
    items, a seed per job."""
    rng = np.random.default_rng(seed)
    kinds = rng.choice(["collatz", "photon", "bfs"], size=n)
    sizes = np.rint(2.0 ** rng.uniform(10.0, 22.0, size=n)).astype(np.int64)
    seeds = rng.integers(1, 1 << 40, size=n)
    return [(str(k), int(s), int(sd)) for k, s, sd in zip(kinds, sizes, seeds)]


def serving_sweep(hz, jobs: int = SERVE_JOBS, loads=(0.2, 0.4, 0.6, 0.8, 1.0)) -> dict:
    """Peak throughput with every job arriving at once, then open-loop Poisson
    arrivals of the same job mix at fractions of that peak: p50/p99 latency
    (arrival to completion seen by the server) and achieved throughput at each
    load, for arrival-order and largest-first admission."""
    out = {
        "workload": f"{jobs} jobs, Collatz / photon walk / BFS with p = 1/3 each, sizes log-uniform in "
                    f"[2^10, 2^22] items, per-job seeds; 3 slots (concurrent streams); the same mix at every load",
    }
    mix = serving_mix(jobs, 1)
    unit = np.random.default_rng(2).exponential(1.0, size=len(mix))  # Poisson gaps at rate 1
    for name, policy in (("fifo", hz.SERVE_FIFO), ("largest_first", hz.SERVE_LARGEST_FIRST)):
        srv = hz.Server(pool_bytes=48 << 30, slots=3, policy=policy)
        try:
            warm = serving_mix(12, 99)
            srv.run_trace(warm, np.zeros(len(warm)))
            res = srv.run_trace(mix, np.zeros(len(mix)))
            span = max(r["end_ms"] for r in res) - min(r["submit_ms"] for r in res)
            peak = len(res) / (span * 1e-3)
            entry = {
                "peak_jobs_per_s": peak,
                "peak_promises_per_s": sum(r["promises"] for r in res) / (span * 1e-3),
                "peak_makespan_ms": span,
                "loads": [],
            }
            for load in loads:
                rate = load * peak  # jobs per second
                arrivals = np.cumsum(unit / rate) * 1e3
                arrivals -= arrivals[0]
                res = srv.run_trace(mix, arrivals)
                lat = np.array([r["end_ms"] - r["submit_ms"] for r in res])
                span = max(r["end_ms"] for r in res) - min(r["submit_ms"] for r in res)
                by_kind = {}
                for kind in ("collatz", "photon", "bfs"):
                    sel = np.array([m[0] == kind for m in mix])
                    by_kind[kind] = {"p50_ms": float(np.percentile(lat[sel], 50)),
                                     "p99_ms": float(np.percentile(lat[sel], 99))}
                entry["loads"].append({
                    "offered_fraction_of_peak": load,
                    "offered_jobs_per_s": rate,
                    "achieved_jobs_per_s": len(res) / (span * 1e-3),
                    "achieved_promises_per_s": sum(r["promises"] for r in res) / (span * 1e-3),
                    "p50_ms": float(np.percentile(lat, 50)),
                    "p99_ms": float(np.percentile(lat, 99)),
                    "max_ms": float(lat.max()),
                    "by_kind": by_kind,
                })
            out[name] = entry
        finally:
            srv.close()
        progress(f"serving mix, {name}")
    return out


def resident_sweep(hz, kind: str = "collatz", jobs: int = SERVE_JOBS, loads=(0.2, 0.4, 0.6, 0.8, 1.0)) -> dict:
    """Jobs of one kind (Collatz: sizes log-uniform in [2^10, 2^22] starts
    below 2^32; photon: the same sizes in photons, default walk parameters;
    per-job seeds), served two ways on the same trace: admitted into a
    running resident kernel (hz_resident_*, arrival order and smallest job
    first) and by the slot server (one launch per job or chunk on 3 streams).
    Peak throughput with every job arriving at once, then Poisson arrivals at
    fractions of each server's own peak; also one 2^22-item job alone."""
    rng = np.random.default_rng(1)
    sizes = np.rint(2.0 ** rng.uniform(10.0, 22.0, size=jobs)).astype(np.int64)
    seeds = rng.integers(1, 1 << 40, size=jobs)
    trace = list(zip(sizes.tolist(), seeds.tolist()))
    unit = np.random.default_rng(2).exponential(1.0, size=jobs)
    items = "starts uniform in [1, 2^32)" if kind == "collatz" else "photons, default walk parameters"
    out = {"workload": f"{jobs} {kind} jobs, sizes log-uniform in [2^10, 2^22] {items}, per-job seeds; "
                       f"the same trace for every server and load"}

    def sweep(srv, run, label):
        run(srv, [(1024, 5)] * 8, np.zeros(8))
        lone = [run(srv, [(1 << 22, 77 + k)], np.zeros(1))[0] for k in range(3)]
        progress(f"{label}, lone jobs")
        res = run(srv, trace, np.zeros(jobs))
        span = max(r["end_ms"] for r in res) - min(r["submit_ms"] for r in res)
        peak = jobs / (span * 1e-3)
        entry = {"lone_2^22_ms": float(np.median([r["end_ms"] - r["submit_ms"] for r in lone])),
                 "peak_jobs_per_s": peak, "peak_makespan_ms": span, "loads": []}
        for load in loads:
            arrivals = np.cumsum(unit / (load * peak)) * 1e3
            arrivals -= arrivals[0]
            res = run(srv, trace, arrivals)
            lat = np.array([r["end_ms"] - r["submit_ms"] for r in res])
            span = max(r["end_ms"] for r in res) - min(r["submit_ms"] for r in res)
            entry["loads"].append({"offered_fraction_of_peak": load, "offered_jobs_per_s": load * peak,
                                   "achieved_jobs_per_s": jobs / (span * 1e-3),
                                   "p50_ms": float(np.percentile(lat, 50)),
                                   "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max())})
            progress(f"{label}, load {load}")
        return entry

    def run_resident(srv, tr, arrivals):
        return srv.run_trace(tr, arrivals, timeout_ms=120000.0)

    def run_slots(srv, tr, arrivals):
        return srv.run_trace([(kind, s, sd) for s, sd in tr], arrivals)

    servers = [("resident_arrival_order", lambda: hz.Resident(pool_bytes=24 << 30, kind=kind), run_resident)]
    if kind == "collatz":  # photon servers claim in arrival order only
        servers.append(("resident_smallest_first",
                        lambda: hz.Resident(pool_bytes=24 << 30, policy=hz.RESIDENT_SMALLEST_FIRST, kind=kind),
                        run_resident))
    servers.append(("slots", lambda: hz.Server(pool_bytes=24 << 30, slots=3), run_slots))
    for name, make, run in servers:
        srv = make()
        try:
            out[name] = sweep(srv, run, f"{kind} serving, {name}")
            if name.startswith("resident"):
                info = srv.info()
                out[name]["launches"] = info["launches"]
                out[name]["jobs_joined_running_launch"] = info["joined_running"]
                out[name]["jobs"] = info["jobs"]
        finally:
            srv.close()
        progress(f"{kind} serving, {name}")
    return out


def _leg_collatz(kind):
    def run(hz):
        starts = hz.collatz_make_starts(SEED, BITS, COUNT)
        steps_out = hz.DeviceArray(COUNT, np.uint32)
        cfg = dict(arena=ARENA_PROMISES // LINK_PROMISES, queues=QUEUES) if kind == "async" else dict(max_launches=1 << 20)
        return hz.collatz_run(kind, steps_out, starts, COUNT, **cfg)
    return run


def _leg_photon(kind):
    def run(hz):
        cfg = {} if kind == "async" else dict(groups=2048, cycles=8)
        return hz.photon_run(kind, hz.photon_params(), max_launches=1 << 20, **cfg)[2]
    return run


def _leg_copy(hz):
    """1 GiB device-to-device copy: the calibration of the HBM byte counters."""
    nbytes = 1 << 30
    a = hz.DeviceArray(nbytes // 4, np.uint32)
    b = hz.DeviceArray(nbytes // 4, np.uint32)
    hz.device_copy(b, a, nbytes)
    hz.check(hz.lib().hz_device_synchronize(), "sync")
    return {"bytes": nbytes}


def _leg_bfs(kind):
    def run(hz):
        batch = hz.BfsBatch.rmat(BFS_SEED, hz.bfs_batch_sizes(BFS_SEED, BFS_GRAPHS, 14.0, 22.0))
        try:
            # Plain stream launches, so that counter collection sees ordinary dispatches.
            return batch.run(kind, max_launches=1 << 16)
        finally:
            batch.close()
    return run


def _leg_barrier(hz):
    counter, hits, violations, st = hz.barrier_run(BARRIER_DEP, BARRIER_NON_DEP, 1024, arena=1 << 17,
                                                   max_launches=1000)
    if counter != 2 * BARRIER_DEP + BARRIER_NON_DEP or violations != 0:
        raise RuntimeError("barrier demo: wrong result")
    return st


def _leg_futures(kind):
    def run(hz):
        starts = hz.DeviceArray.from_numpy(np.arange(1, FUTURES_LEAVES + 1, dtype=np.uint64))
        return hz.futures_run(kind, starts, FUTURES_READERS if kind == "async" else 0)[3]
    return run


LEGS = {
    "barrier_async": _leg_barrier,
    "bfs_async": _leg_bfs("async"),
    "bfs_event": _leg_bfs("event"),
    "collatz_async": _leg_collatz("async"),
    "collatz_event": _leg_collatz("event"),
    "photon_async": _leg_photon("async"),
    "photon_event": _leg_photon("event"),
    "copy": _leg_copy,
    "futures_async": _leg_futures("async"),
    "futures_event": _leg_futures("event"),
}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--groups", type=int, default=0, help="AsyncProgram groups (0 = fill the device)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-event", action="store_true", help="skip the EventProgram comparator")
    ap.add_argument("--no-photon", action="store_true", help="skip the photon-walk Async vs Event comparison")
    ap.add_argument("--no-bfs", action="store_true", help="skip the batched BFS Async vs Event comparison")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving load sweep")
    ap.add_argument("--no-futures", action="store_true", help="skip the value-returning futures tree")
    ap.add_argument("--no-barrier", action="store_true", help="skip the remapping-barrier demo")
    ap.add_argument("--leg", choices=sorted(LEGS), help="run one leg once (for counter collection) and exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this runtime is measured on one GPU; --gpus must be 1 (sharding a program "
                                   "across GPUs is not built)"}))
        return 2

    import harmonize_amd as hz

    if args.leg:
        print(json.dumps({"leg": args.leg, **LEGS[args.leg](hz)}), flush=True)
        return 0

    facts = hz.device_facts()
    starts = hz.collatz_make_starts(SEED, BITS, COUNT)
    steps_out = hz.DeviceArray(COUNT, np.uint32)
    arena_links = ARENA_PROMISES // LINK_PROMISES
    cfg = dict(arena=arena_links, groups=args.groups, queues=QUEUES)

    for _ in range(args.warmup):
        hz.collatz_run("async", steps_out, starts, COUNT, **cfg)
    hz.check(hz.lib().hz_device_synchronize(), "sync")

    runs = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        runs.append(hz.collatz_run("async", steps_out, starts, COUNT, **cfg))
    hz.check(hz.lib().hz_device_synchronize(), "sync")
    wall = time.perf_counter() - t0
    steps_async = steps_out.numpy()  # compared below with the EventProgram and one-thread-per-start runs

    promises = runs[-1]["promises"]
    total_promises = sum(r["promises"] for r in runs)
    value = total_promises / wall
    kernel_ms = float(np.mean([r["elapsed_ms"] for r in runs]))
    launches = int(np.mean([r["launches"] for r in runs]))
    per_launch_ms = kernel_ms / max(1, launches)
    per_launch_s = per_launch_ms * 1e-3
    cus = facts["compute_units"]
    # The bound the headline kernel reaches: scalar instruction issue. A CU's
    # four SIMDs share one scalar unit that issues at most one instruction per
    # clock, so the chip's peak is CUs x clock.
    salu_peak = cus * CLOCK_GHZ  # G SALU instructions/s
    salu_achieved = ISSUE_COUNTS["salu"] / max(1, launches) / per_launch_s / 1e9
    # Beside it: the LDS bytes the stash moves (each promise that passes
    # through it is written into a link and read back out, 2 x 16 B; promises
    # run by link forwarding stay in registers) against 128 B per CU per
    # clock, and the measured HBM traffic against the HBM peak.
    stashed = promises - runs[-1]["forwarded"]
    lds_achieved = 2 * PROMISE_BYTES * stashed / max(1, launches) / per_launch_s / 1e9
    lds_peak = cus * LDS_BYTES_PER_CLOCK_PER_CU * CLOCK_GHZ
    hbm_achieved = measured_traffic_bytes() / per_launch_s / 1e9
    alg_equiv = algorithmic_bytes(promises, COUNT) / max(1, launches) / per_launch_s / 1e9

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
            "forwarded_per_step": runs[-1]["forwarded"],
            "launches_per_step": launches,
            "spills": runs[-1]["spills"],
            "fills": runs[-1]["fills"],
            "device": facts["arch"],
            "compute_units": facts["compute_units"],
        },
        "roofline": {
            "bound": "salu",
            "achieved": salu_achieved,
            "peak": salu_peak,
            "unit": "G SALU instructions/s",
            "frac": salu_achieved / salu_peak,
            "traffic": measured_traffic_bytes(),
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0)",
            "kernel": "_hz_exec_kernel<AsyncProgram<CollatzSpec>>",
            "kernel_ms": per_launch_ms,
            "peak_model": f"one scalar instruction per CU per clock: {cus} CUs x {CLOCK_GHZ} GHz",
            "counts_per_launch": {k: v / max(1, launches) for k, v in ISSUE_COUNTS.items()},
            "counters": "profiles/collatz_async_issue_counters.md",
            "lds": {"achieved": lds_achieved, "peak": lds_peak, "unit": "GB/s", "frac": lds_achieved / lds_peak,
                    "rule": "2 x 16 B per promise through the stash (write + read; forwarded promises skip "
                            "it); peak 128 B per CU per clock"},
            "hbm": {"achieved": hbm_achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": hbm_achieved / HBM_PEAK_GBPS, "rule": "measured traffic per launch / launch time"},
            "algorithmic_hbm_equivalent": {
                "achieved": alg_equiv, "unit": "GB/s",
                "rule": "2 x 16 B per promise + 12 B per start: the bytes an HBM-queued runtime would move "
                        "for the same promises (informational; the stash keeps them in LDS)",
            },
        },
    }

    progress("Collatz AsyncProgram")
    if not args.no_event:
        ev = hz.collatz_run("event", steps_out, starts, COUNT, max_launches=1 << 20)
        if not np.array_equal(steps_out.numpy(), steps_async):
            raise RuntimeError("collatz: AsyncProgram and EventProgram step counts differ")
        out["event"] = {
            "value": ev["promises"] / (ev["elapsed_ms"] * 1e-3),
            "unit": "promises/s",
            "ms": ev["elapsed_ms"],
            "launches": ev["launches"],
            "groups": ev["groups"],
        }
        out["async_vs_event"] = value / out["event"]["value"]
        progress("Collatz EventProgram")

    if not args.no_photon:
        out["photon"] = photon_comparison(hz)
        progress("photon walk")

    if not args.no_bfs:
        out["bfs"] = bfs_comparison(hz)
        progress("BFS batch")

    if not args.no_futures:
        out["futures"] = futures_comparison(hz)
        progress("futures tree")
    if not args.no_barrier:
        out["barrier"] = barrier_demo(hz)
        progress("barrier demo")

    if not args.no_serve:
        out["serving"] = serving_sweep(hz)
        out["serving_resident"] = resident_sweep(hz)
        out["serving_resident_photon"] = resident_sweep(hz, "photon")

    pl = hz.collatz_run("plain", steps_out, starts, COUNT)
    if not np.array_equal(steps_out.numpy(), steps_async):
        raise RuntimeError("collatz: AsyncProgram and one-thread-per-start step counts differ")
    out["plain_kernel"] = {
        "ms": pl["elapsed_ms"],
        "steps_per_s": promises / (pl["elapsed_ms"] * 1e-3),
        "note": "one thread per start, no runtime (steps computed per second, same work)",
    }

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        progress("CPU baseline")

    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
