"""bench.py's contract that runs without a GPU: the module imports, every
profiling leg is defined, and --gpus above 1 exits with a message."""

import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_bench_imports_and_defines_legs():
    sys.path.insert(0, str(ROOT))
    import bench

    assert set(bench.LEGS) == {"barrier_async", "bfs_async", "bfs_event", "collatz_async", "collatz_event",
                               "photon_async", "photon_event", "copy", "futures_async", "futures_event"}
    assert bench.measured_traffic_bytes() > 0
    mix = bench.serving_mix(1000, 1)
    kinds = {k for k, _, _ in mix}
    sizes = [s for _, s, _ in mix]
    assert kinds == {"collatz", "photon", "bfs"}
    assert min(sizes) >= 1 << 10 and max(sizes) <= 1 << 22


def test_more_than_one_gpu_exits_with_a_message():
    proc = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                          timeout=120)
    assert proc.returncode == 2
    assert "error" in json.loads(proc.stdout.strip().splitlines()[-1])
