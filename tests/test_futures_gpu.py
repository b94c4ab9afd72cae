"""Value-returning operations and futures on the GPU runtime, compared with
the oracle's host executors on the same starts (bit-exact totals, echoes,
forwarded values and promise counts), with the closed-form Collatz count
(collaz_hrm.cpp:157-169) where the starts are 1..n.

Paths covered: a promise returned from eval taking over the caller's return
address (Leaf, Steps), a plain value resolving a UnitBarrier future (Join), a
RemappingBarrier future holding awaiting promises (Echo) and ReturnOp
forwarding a future's value (Forward); both the held path (readers claimed
before the total resolves) and the already-released path."""

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def _run(kind, starts, readers, **cfg):
    import harmonize_amd as hz

    dev = hz.DeviceArray.from_numpy(np.ascontiguousarray(starts, dtype=np.uint64))
    return hz.futures_run(kind, dev, readers, **cfg)


def _check_against_oracle(starts, readers, total, echo, forwarded, st):
    ref_total, ref_echo, ref_forwarded, ref_st = O.futures_program(starts, readers, threads=0)
    assert total == ref_total
    assert np.array_equal(echo, ref_echo)
    assert np.array_equal(forwarded, ref_forwarded)
    assert st["promises"] == ref_st.promises
    assert np.array_equal(echo, total + np.arange(readers, dtype=np.uint64))
    assert np.all(forwarded == total)


@pytest.mark.parametrize("leaves,readers", [(1, 0), (1, 5), (2, 1), (1 << 10, 100), (1 << 16, 4096)])
def test_async_sequential_starts(leaves, readers):
    starts = np.arange(1, leaves + 1, dtype=np.uint64)
    total, echo, forwarded, st = _run("async", starts, readers)
    _check_against_oracle(starts, readers, total, echo, forwarded, st)
    assert st["forwarded"] == 0  # the futures specification turns link forwarding off


def test_async_closed_form_2p20():
    starts = np.arange(1, (1 << 20) + 1, dtype=np.uint64)
    total, echo, forwarded, st = _run("async", starts, 1 << 14)
    assert total == 139_348_407 - (1 << 20)
    assert st["promises"] == (1 << 20) + 139_348_407 + (1 << 20) - 1 + 3 * (1 << 14)
    assert np.array_equal(echo, total + np.arange(1 << 14, dtype=np.uint64))
    assert np.all(forwarded == total)


def test_async_random_starts():
    rng = np.random.default_rng(11)
    starts = rng.integers(1, 1 << 32, size=1 << 18, dtype=np.uint64)
    total, echo, forwarded, st = _run("async", starts, 20000)
    _check_against_oracle(starts, 20000, total, echo, forwarded, st)


def test_async_relaunches_and_small_grid():
    # A small cycle budget ends launches early: held Echo / ReturnOp links and
    # stash links must survive the relaunches. (A group runs up to eight links
    # per cycle, so the budget is far below what one launch needs.)
    starts = np.arange(1, (1 << 14) + 1, dtype=np.uint64)
    total, echo, forwarded, st = _run("async", starts, 3000, groups=64, cycles=32, max_launches=100000,
                                      flags=2)  # harmonize_amd.RUN_AUDIT
    assert st["launches"] > 1 and st["audits"] == st["launches"]
    _check_against_oracle(starts, 3000, total, echo, forwarded, st)


@pytest.mark.parametrize("leaves", [1, 1 << 8, 1 << 16])
def test_event_tree(leaves):
    starts = np.arange(1, leaves + 1, dtype=np.uint64)
    total, _, _, st = _run("event", starts, 0)
    ref_total, _, _, ref_st = O.futures_program(starts, 0, threads=0)
    assert total == ref_total
    assert st["promises"] == ref_st.promises


def test_event_rejects_readers():
    import harmonize_amd as hz

    with pytest.raises(hz.HarmonizeError, match="EventProgram"):
        _run("event", np.arange(1, 5, dtype=np.uint64), 2)


def test_rejects_non_power_of_two():
    import harmonize_amd as hz

    with pytest.raises(hz.HarmonizeError, match="power of two"):
        _run("async", np.arange(1, 4, dtype=np.uint64), 0)
