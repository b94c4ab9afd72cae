"""GPU parity of the Collatz programs against the oracle: bit-exact step counts
and exact promise counts, through the C ABI."""

import ctypes

import numpy as np
import pytest

import harmonize_amd as hz
import oracle_lib as O

pytestmark = pytest.mark.gpu


def run(kind, n, starts=None, **cfg):
    steps = hz.DeviceArray(n, np.uint32)
    st = hz.collatz_run(kind, steps, starts, n, **cfg)
    return steps.numpy(), st


@pytest.mark.parametrize("kind", ["plain", "event", "async"])
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 1000, 4097])
def test_small_and_ragged(kind, n):
    ref, promises = O.collatz_steps(None, count=n)
    got, st = run(kind, n, max_launches=100000)
    assert np.array_equal(got, ref)
    if kind != "plain":
        assert st["promises"] == promises


@pytest.mark.parametrize("kind", ["event", "async"])
def test_zero_starts(kind):
    got, st = run(kind, 0)
    assert got.size == 0 and st["promises"] == 0


@pytest.mark.parametrize("kind", ["plain", "event", "async"])
def test_c0_index_starts_2pow20(kind):
    n = 1 << 20
    ref, promises = O.collatz_steps(None, count=n)
    got, st = run(kind, n, max_launches=100000)
    assert np.array_equal(got, ref)
    if kind != "plain":
        assert st["promises"] == promises == 139348387


def test_starts_generator_matches_oracle():
    n = 1 << 20
    dev = hz.collatz_make_starts(0x5EED, 40, n)
    assert np.array_equal(dev.numpy(), O.collatz_make_starts(0x5EED, 40, n))


@pytest.fixture(scope="module")
def c1_full():
    n = 1 << 26
    host = O.collatz_make_starts(0x5EED, 40, n)
    ref, promises = O.collatz_steps(host)
    dev = hz.collatz_make_starts(0x5EED, 40, n)
    return n, dev, ref, promises


@pytest.mark.parametrize(
    "kind,cfg",
    [
        ("plain", {}),
        ("async", {}),
        ("event", {}),
        # the configuration the benchmark times: an arena of 2^24 promises
        # (2^18 links) behind 16 link queues
        ("async", {"arena": 1 << 18, "queues": 16}),
    ],
)
def test_c1_full_scale_random_starts(c1_full, kind, cfg):
    n, dev, ref, promises = c1_full
    got, st = run(kind, n, starts=dev, max_launches=100000, **cfg)
    assert np.array_equal(got, ref)
    assert int((got == hz.COLLATZ_OVERFLOW).sum()) == 3
    if kind != "plain":
        assert st["promises"] == promises


def test_async_relaunch_and_spill_path():
    # A small cycle budget ends every launch early: each group spills its whole
    # stash to the arena and the next launch refills it.
    n = 1 << 22
    host = O.collatz_make_starts(11, 40, n)
    ref, promises = O.collatz_steps(host)
    dev = hz.DeviceArray.from_numpy(host)
    got, st = run("async", n, starts=dev, cycles=500, groups=512, arena=1 << 16, max_launches=100000,
                  flags=hz.RUN_AUDIT)
    assert np.array_equal(got, ref)
    assert st["promises"] == promises
    assert st["launches"] > 1 and st["spills"] > 0 and st["fills"] == st["spills"]
    # the host audit of queues, arena and promise accounts after every launch,
    # with link forwarding running promises straight from registers
    assert st["audits"] == st["launches"]
    assert 0 < st["forwarded"] < st["promises"]


def test_async_forwarding_keeps_results_and_counts():
    """Link forwarding: a full link of odd steps is followed by a full link
    of even steps, which runs from registers. Starts 1..2^20: step counts and
    the promise count are the oracle's, and a share of the promises (at most
    the even steps that follow an odd one, so under half of all steps) ran
    forwarded."""
    n = 1 << 20
    host = np.arange(1, n + 1, dtype=np.uint64)
    ref, promises = O.collatz_steps(host)
    dev = hz.DeviceArray.from_numpy(host)
    got, st = run("async", n, starts=dev, max_launches=100000)
    assert np.array_equal(got, ref)
    assert st["promises"] == promises
    assert 0 < st["forwarded"] < st["promises"] // 2


@pytest.mark.parametrize("queues", [1, 3, 16])
def test_async_queue_counts(queues):
    n = 1 << 20
    ref, promises = O.collatz_steps(None, count=n)
    got, st = run("async", n, cycles=300, groups=1024, queues=queues, max_launches=100000)
    assert np.array_equal(got, ref) and st["promises"] == promises


def test_async_arena_exhaustion_is_reported():
    n = 1 << 20
    with pytest.raises(hz.HarmonizeError, match="arena"):
        run("async", n, cycles=50, groups=1024, arena=4, max_launches=1000)


def test_event_small_queue_overflow_is_reported():
    with pytest.raises(hz.HarmonizeError, match="overflow"):
        run("event", 1 << 20, arena=64 * 64, groups=4096, cycles=1, max_launches=100000)


@pytest.mark.parametrize("suffix,size,cycles", [("collatz_async", 1 << 12, 1 << 20), ("collatz_event", 1 << 20, 4)])
def test_reference_shaped_entry_points(suffix, size, cycles):
    L = hz.lib()
    n = 50000
    steps = hz.DeviceArray(n, np.uint32)
    it = hz.DeviceArray.from_numpy(np.array([0, n], dtype=np.uint32))  # AtomicIter {value, limit}
    state = (ctypes.c_uint64 * 3)(0, steps.ptr, it.ptr)
    dev_state = getattr(L, f"alloc_state_{suffix}")()
    assert dev_state, hz.last_error()
    getattr(L, f"store_state_device_{suffix}")(dev_state, ctypes.addressof(state))
    back = (ctypes.c_uint64 * 3)()
    getattr(L, f"load_state_device_{suffix}")(ctypes.addressof(back), dev_state)
    assert list(back) == list(state)
    inst = getattr(L, f"alloc_program_{suffix}")(dev_state, size)
    assert inst, hz.last_error()
    getattr(L, f"init_program_{suffix}")(inst, 256)
    hz.check_void("init_program")
    launches = 0
    while True:
        getattr(L, f"exec_program_{suffix}")(inst, 256, cycles)
        hz.check_void("exec_program")
        launches += 1
        done = getattr(L, f"complete_{suffix}")(inst)
        assert done >= 0, hz.last_error()
        if done:
            break
        assert launches < 100000
    getattr(L, f"clear_flags_{suffix}")(inst)
    getattr(L, f"free_program_{suffix}")(inst)
    getattr(L, f"free_state_{suffix}")(dev_state)
    ref, _ = O.collatz_steps(None, count=n)
    assert np.array_equal(steps.numpy(), ref)


@pytest.mark.gpu
def test_tuner_collatz_small():
    import harmonize_amd as hz

    n = 1 << 16
    starts = hz.collatz_make_starts(0x5EED, 40, n)
    steps = hz.DeviceArray(n, np.uint32)
    res = hz.tune_collatz(steps, starts, n, rounds=1)
    assert res["trials"] == len(res["log"]) and res["best_ms"] > 0
    best = {k: v for k, v in res["best"].items() if v}
    hz.collatz_run("async", steps, starts, n, **best)
    want, _ = O.collatz_steps(starts.numpy())
    assert np.array_equal(steps.numpy(), want)
