"""CPU pinning of the futures specification (value-returning operations,
Future, UnitBarrier, RemappingBarrier futures, ReturnOp) on the oracle's host
executors.

The reference declares Future / ReturnOp / Return (async/future.h:7-60,
async/promise.h:326-434) but leaves their resolution unfinished, so there is no
reference output to match. The host executors' results are pinned by what the
mathematics fixes instead:
  - the total is the sum of the leaves' Collatz step counts: checked against an
    independent numpy restatement, and for starts 1..2^20 against the
    reference's own closed-form count (collaz_hrm.cpp:157-169: the sum of
    steps + 1 over n = 1..2^20 is 139,348,407);
  - every reader sees the total: echo[r] = total + r, forwarded[r] = total;
  - every promise runs exactly once: Leaf (one per leaf), Steps (steps + 1 per
    leaf), Join (one per internal node), and per reader an Echo, a Forward and
    the ReturnOp that forwards the total.
"""
import numpy as np
import pytest

import oracle_lib as O


def steps_numpy(starts):
    v = np.asarray(starts, dtype=np.uint64).copy()
    s = np.zeros(len(v), dtype=np.uint64)
    live = v > 1
    while live.any():
        x = v[live]
        v[live] = np.where(x & np.uint64(1), np.uint64(3) * x + np.uint64(1), x >> np.uint64(1))
        s[live] += np.uint64(1)
        live = v > 1
    return s


def expected_promises(steps, readers):
    leaves = len(steps)
    return leaves + int(steps.sum()) + leaves + (leaves - 1) + 3 * readers


@pytest.mark.parametrize("threads", [1, 4])
def test_tree_total_and_readers(threads):
    starts = np.arange(1, (1 << 12) + 1, dtype=np.uint64)
    steps = steps_numpy(starts)
    total, echo, forwarded, st = O.futures_program(starts, readers=300, threads=threads)
    assert total == int(steps.sum())
    assert np.array_equal(echo, total + np.arange(300, dtype=np.uint64))
    assert np.all(forwarded == total)
    assert st.promises == expected_promises(steps, 300)


def test_closed_form_total_2p20():
    starts = np.arange(1, (1 << 20) + 1, dtype=np.uint64)
    total, _, _, st = O.futures_program(starts, readers=0, threads=0)
    # collaz_hrm.cpp's checker: the sum of (steps + 1) over 1..2^20 is 139,348,407.
    assert total == 139_348_407 - (1 << 20)
    assert st.promises == (1 << 20) + 139_348_407 + (1 << 20) - 1


def test_random_starts_match_numpy():
    rng = np.random.default_rng(7)
    starts = rng.integers(1, 1 << 32, size=1 << 10, dtype=np.uint64)
    steps = steps_numpy(starts)
    total, echo, forwarded, st = O.futures_program(starts, readers=64, threads=4)
    assert total == int(steps.sum())
    assert np.all(forwarded == total) and np.array_equal(echo, total + np.arange(64, dtype=np.uint64))
    assert st.promises == expected_promises(steps, 64)


def test_single_leaf_resolves_the_root_directly():
    total, echo, forwarded, st = O.futures_program(np.array([27], dtype=np.uint64), readers=3)
    assert total == 111  # steps(27), the classic value
    assert list(echo) == [111, 112, 113] and list(forwarded) == [111] * 3
    assert st.promises == 1 + 112 + 0 + 9


def test_trivial_starts():
    # steps(0) = steps(1) = 0: each chain ends at once (v <= 1 stops).
    total, _, _, st = O.futures_program(np.array([0, 1, 1, 0], dtype=np.uint64), readers=0)
    assert total == 0
    assert st.promises == 4 + 4 + 3


def test_leaf_count_must_be_a_power_of_two():
    with pytest.raises(ValueError):
        O.futures_program(np.arange(1, 4, dtype=np.uint64), readers=0)


def test_library_checks_futures_arguments_before_any_device_work():
    # hz_futures_run validates its arguments before touching the GPU, so the
    # checks are exercised here without one.
    import ctypes

    import harmonize_amd as hz

    lib = hz.lib()
    total = ctypes.c_uint64(0)
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: every case fails its checks first

    def run(kind, leaves, readers, starts=fake, echo=fake, fwd=fake, out=ctypes.byref(total)):
        rc = lib.hz_futures_run(kind, starts, leaves, readers, echo, fwd, out, None, None)
        msg = lib.hz_last_error()
        return rc, (msg.decode() if msg else "")

    rc, msg = run(hz.ASYNC, 3, 0)
    assert rc != 0 and "power of two" in msg
    rc, msg = run(hz.ASYNC, 0, 0)
    assert rc != 0 and "power of two" in msg
    rc, msg = run(hz.ASYNC, 1 << 27, 0)
    assert rc != 0 and "power of two" in msg
    rc, msg = run(hz.ASYNC, 4, (1 << 24) + 1)
    assert rc != 0 and "readers" in msg
    rc, msg = run(hz.ASYNC, 4, 2, echo=None)
    assert rc != 0 and "null" in msg
    rc, msg = run(hz.EVENT, 4, 2)
    assert rc != 0 and "EventProgram" in msg
    rc, msg = run(7, 4, 0)
    assert rc != 0 and "kind" in msg
