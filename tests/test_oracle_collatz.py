"""Pins the Collatz oracle (oracle/oracle.cpp and the Collatz specification run
on the host executors) against known answers, an independent numpy
restatement and Python big-integer brute force."""

import json
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = json.loads((Path(__file__).parent / "golden" / "collatz_known.json").read_text())
OVERFLOW = 0xFFFFFFFF


def numpy_steps(starts):
    """Vectorised Collatz step counts (values must stay below 2^63)."""
    v = np.array(starts, dtype=np.uint64)
    steps = np.zeros(v.size, dtype=np.uint32)
    peak = v.copy()
    live = v > 1
    while live.any():
        odd = (v & np.uint64(1)) == 1
        v = np.where(live & odd, v * np.uint64(3) + np.uint64(1), np.where(live, v >> np.uint64(1), v))
        peak = np.maximum(peak, v)
        steps += live.astype(np.uint32)
        live = v > 1
    return steps, peak


def bigint_steps(start):
    v, steps, evals = start, 0, 1
    while v > 1:
        if v % 2 == 0:
            v //= 2
        else:
            v = 3 * v + 1
            if v >= 1 << 64:
                return OVERFLOW, steps + 1
        steps += 1
    return steps, steps + 1


def test_known_single_values():
    for n, s in GOLDEN["steps"].items():
        assert O.lib().oracle_collatz_steps_one(int(n)) == s, n


def test_range_1_to_2pow20_known_totals():
    starts = np.arange(1, (1 << 20) + 1, dtype=np.uint64)
    steps, promises = O.collatz_steps(starts)
    g = GOLDEN["range_1_to_2pow20"]
    assert promises == g["promises"]
    assert int(steps.sum(dtype=np.uint64)) == g["sum_steps"]
    assert int(steps.max()) == g["max_steps"]
    assert int(starts[int(steps.argmax())]) == g["argmax"]


def test_numpy_restatement_agrees_and_peak_fits_64_bits():
    starts = np.arange(0, 1 << 20, dtype=np.uint64)
    ref, peak = numpy_steps(starts)
    got, _ = O.collatz_steps(None, count=1 << 20)
    assert np.array_equal(got, ref)
    assert int(peak.max()) == GOLDEN["range_1_to_2pow20"]["max_trajectory_value"]


def test_overflow_policy_matches_bigint():
    for s in GOLDEN["overflow_starts_below_2pow40"]:
        assert bigint_steps(s)[0] == OVERFLOW
        assert O.lib().oracle_collatz_steps_one(s) == OVERFLOW
    # Promise counts of overflowing starts: evaluations up to the failing 3v+1.
    starts = np.array(GOLDEN["overflow_starts_below_2pow40"] + [27, 1, 0], dtype=np.uint64)
    steps, promises = O.collatz_steps(starts)
    assert promises == sum(bigint_steps(int(s))[1] for s in starts)


def test_random_large_starts_against_bigint():
    starts = O.collatz_make_starts(7, 40, 2000)
    steps, promises = O.collatz_steps(starts)
    expect = [bigint_steps(int(s)) for s in starts]
    assert [int(x) for x in steps] == [e[0] for e in expect]
    assert promises == sum(e[1] for e in expect)


def test_make_starts_range_and_determinism():
    a = O.collatz_make_starts(0x5EED, 40, 1 << 16)
    b = O.collatz_make_starts(0x5EED, 40, 1 << 16)
    c = O.collatz_make_starts(0x5EEE, 40, 1 << 16)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 1 and a.max() < (1 << 40)
    # Uniform enough: mean close to 2^39.
    assert abs(float(a.mean()) / float(1 << 39) - 1.0) < 0.02


@pytest.mark.parametrize("threads", [1, 0])
def test_spec_on_host_executors_matches_closed_form(threads):
    n = 1 << 16
    starts = O.collatz_make_starts(3, 40, n)
    ref, promises = O.collatz_steps(starts)
    got, st = O.collatz_program(starts, threads=threads)
    assert np.array_equal(got, ref)
    assert st.promises == promises
    got2, st2 = O.collatz_program(None, count=n, threads=threads)
    ref2, p2 = O.collatz_steps(None, count=n)
    assert np.array_equal(got2, ref2) and st2.promises == p2


def test_host_executor_handles_empty_and_tiny():
    for n in (0, 1, 2, 63, 64, 65):
        got, st = O.collatz_program(None, count=n, threads=1)
        ref, p = O.collatz_steps(None, count=n)
        assert np.array_equal(got, ref) and st.promises == p
