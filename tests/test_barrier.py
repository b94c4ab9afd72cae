"""RemappingBarrier: the reference barrier demo's expected count
(dep * 2 + non_dep, examples/cpp/barrier_demo/demo.cpp:214), every held
promise run exactly once, and none before its barrier's release — on the host
executors (CPU) and on the GPU runtime."""

import numpy as np
import pytest

import oracle_lib as O

CASES = [
    (1000, 500, 7),
    (20000, 3000, 64),
    (5000, 0, 1),        # one barrier: every await contends on one partial link
    (0, 1000, 3),        # nothing waits
    (64 * 64 + 1, 10, 2),  # crosses many link boundaries
    (10, 10, 100),       # more barriers than dependents
]


@pytest.mark.parametrize("threads", [1, 0])
@pytest.mark.parametrize("dep,non_dep,barriers", CASES)
def test_host_executors(dep, non_dep, barriers, threads):
    counter, hits, violations, st = O.barrier_program(dep, non_dep, barriers, threads=threads)
    assert counter == 2 * dep + non_dep
    assert (hits == 1).all()
    assert violations == 0
    assert st.promises == dep  # held steps run as promises; the others are evaluated in make_work


@pytest.mark.gpu
@pytest.mark.parametrize("dep,non_dep,barriers", CASES + [(200000, 50000, 1000)])
def test_gpu_async(dep, non_dep, barriers):
    import harmonize_amd as hz

    counter, hits, violations, st = hz.barrier_run(dep, non_dep, barriers, max_launches=1000)
    ref_counter, ref_hits, ref_violations, _ = O.barrier_program(dep, non_dep, barriers, threads=0)
    assert counter == ref_counter == 2 * dep + non_dep
    assert np.array_equal(hits, ref_hits)
    assert violations == 0
    assert st["promises"] == dep


@pytest.mark.gpu
def test_gpu_held_links_survive_relaunch():
    import harmonize_amd as hz

    dep, non_dep, barriers = 100000, 1000, 16
    counter, hits, violations, st = hz.barrier_run(dep, non_dep, barriers, cycles=4, groups=64,
                                                   max_launches=100000, flags=hz.RUN_AUDIT)
    assert counter == 2 * dep + non_dep
    assert st["launches"] > 1 and st["audits"] == st["launches"]
    assert (hits == 1).all() and violations == 0
