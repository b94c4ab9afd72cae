"""The allocation-free Collatz entry point (hz_collatz_run_noalloc) against
the allocating one (hz_collatz_run, HZ_ASYNC) through the C ABI: identical
step counts (bit-exact) and promise counts on two calls in a row, the second
with a smaller input, so that whatever the first call left in the shared
device array cannot leak into the second. The oracle checks both."""

import ctypes

import numpy as np
import pytest

import harmonize_amd as hz
import oracle_lib as O

pytestmark = pytest.mark.gpu


def run_both(host_starts, **cfg):
    n = host_starts.size
    starts = hz.DeviceArray.from_numpy(host_starts)
    steps_a = hz.DeviceArray(n, np.uint32)
    st_a = hz.collatz_run("async", steps_a, starts, n, max_launches=100000, **cfg)
    steps_b = hz.DeviceArray(n, np.uint32)
    st_b = hz.collatz_run_noalloc(steps_b, starts, n, max_launches=100000, **cfg)
    return steps_a.numpy(), st_a, steps_b.numpy(), st_b


def test_noalloc_matches_allocating_entry_point_twice_shrinking():
    for n, seed in ((1 << 20, 0x5EED), (12_345, 0xBEEF)):
        host = O.collatz_make_starts(seed, 40, n)
        ref, promises = O.collatz_steps(host)
        got_a, st_a, got_b, st_b = run_both(host)
        print(f"noalloc n={n} promises={st_b['promises']} launches={st_b['launches']}", flush=True)
        assert np.array_equal(got_b, got_a)
        assert np.array_equal(got_b, ref)
        assert st_b["promises"] == st_a["promises"] == promises


def test_noalloc_small_arena_and_many_launches():
    # A short cycle budget ends launches early, so the stash goes to the
    # arena between launches and the next launch resumes it from there.
    n = 1 << 12
    host = O.collatz_make_starts(7, 40, n)
    ref, promises = O.collatz_steps(host)
    steps = hz.DeviceArray(n, np.uint32)
    starts = hz.DeviceArray.from_numpy(host)
    st = hz.collatz_run_noalloc(steps, starts, n, groups=64, cycles=64, arena=4096, queues=16,
                                max_launches=100000)
    assert st["launches"] > 1 and st["spills"] > 0
    assert np.array_equal(steps.numpy(), ref)
    assert st["promises"] == promises


def test_noalloc_refuses_sizes_beyond_the_static_array():
    steps = hz.DeviceArray(16, np.uint32)
    with pytest.raises(hz.HarmonizeError, match="static maximum"):
        hz.collatz_run_noalloc(steps, None, 16, arena=16385)
    with pytest.raises(hz.HarmonizeError, match="static maximum"):
        hz.collatz_run_noalloc(steps, None, 16, queues=17)
