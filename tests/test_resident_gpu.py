"""Resident serving on the GPU: Collatz jobs admitted into a running
persistent kernel through the doorbell ring. Every job's starts and step
counts against the oracle; more jobs than the job table holds, and more than
the ring holds, published while one launch runs; the smallest-first claim
order; a launch that ends when idle and the next submission launching again;
launches cut short by their lifetime, carried on by the next launch; bad
requests refused; jobs waiting for item space, in order, and released
outputs freeing it."""

import ctypes
import time

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

BITS = 32
DRAIN_MS = 60000.0


@pytest.fixture(scope="module")
def hz():
    import harmonize_amd

    return harmonize_amd


def check_job(res, i, size, seed):
    starts = O.collatz_make_starts(seed, BITS, size)
    assert np.array_equal(res.output(i, 1), starts), (i, size, seed)
    want, _ = O.collatz_steps(starts)
    assert np.array_equal(res.output(i, 0), want), (i, size, seed)


@pytest.mark.parametrize("size", [1, 5000, (1 << 20) + 7, 1 << 22])
def test_job_matches_oracle(hz, size):
    res = hz.Resident(keep_outputs=True)
    try:
        seed = 0xA11CE + size
        i = res.submit(size, seed)
        res.drain(DRAIN_MS)
        assert res.job(i)["status"] == hz.JOB_DONE
        check_job(res, i, size, seed)
        info = res.info()
        assert info["running"] == 0 and info["in_flight"] == 0 and info["launches"] >= 1
    finally:
        res.close()


def test_many_jobs_join_one_launch(hz):
    """300 jobs submitted back to back: more than the 64-entry job table and
    the 256-entry ring, so entries and ring slots are reused while the launch
    runs; each job's outputs are its own."""
    rng = np.random.default_rng(7)
    sizes = np.rint(2.0 ** rng.uniform(4, 14, size=300)).astype(int)
    seeds = rng.integers(1, 1 << 40, size=300)
    res = hz.Resident(keep_outputs=True)
    try:
        ids = [res.submit(int(s), int(sd)) for s, sd in zip(sizes, seeds)]
        res.drain(DRAIN_MS)
        info = res.info()
        assert info["joined_running"] > 0
        assert info["launches"] < len(ids)
        for i, s, sd in zip(ids, sizes, seeds):
            assert res.job(i)["status"] == hz.JOB_DONE
            check_job(res, i, int(s), int(sd))
    finally:
        res.close()


def test_smallest_first_lets_small_jobs_overtake(hz):
    """A 2^22-start job followed by small ones: claimed smallest first, every
    small job finishes before the large one, and all outputs are right."""
    res = hz.Resident(keep_outputs=True, policy=hz.RESIDENT_SMALLEST_FIRST)
    try:
        big = res.submit(1 << 22, 99)
        small = [res.submit(1000 + 37 * k, 1000 + k) for k in range(20)]
        res.drain(DRAIN_MS)
        big_end = res.job(big)["end_ms"]
        assert all(res.job(i)["end_ms"] < big_end for i in small)
        check_job(res, big, 1 << 22, 99)
        for k, i in enumerate(small):
            check_job(res, i, 1000 + 37 * k, 1000 + k)
    finally:
        res.close()


def test_idle_launch_ends_and_next_submission_relaunches(hz):
    res = hz.Resident(keep_outputs=True, idle_us=200)
    try:
        a = res.submit(3000, 5)
        res.drain(DRAIN_MS)
        first = res.info()
        assert first["running"] == 0
        b = res.submit(4000, 6)
        res.drain(DRAIN_MS)
        second = res.info()
        assert second["launches"] == first["launches"] + 1
        check_job(res, a, 3000, 5)
        check_job(res, b, 4000, 6)
    finally:
        res.close()


def test_short_lifetime_launches_carry_on(hz):
    """With a 1 ms launch lifetime a 2^22-start job spans several launches:
    the job table and counters carry over, and the result is unchanged."""
    res = hz.Resident(keep_outputs=True, life_ms=1)
    try:
        i = res.submit(1 << 22, 1234)
        j = res.submit(777, 4321)
        res.drain(DRAIN_MS)
        assert res.info()["launches"] > 1
        check_job(res, i, 1 << 22, 1234)
        check_job(res, j, 777, 4321)
    finally:
        res.close()


def test_open_loop_trace(hz):
    """A Poisson trace of 40 jobs (sizes log-uniform in [2^10, 2^18]):
    outputs equal the oracle's and latencies are measured from arrival."""
    rng = np.random.default_rng(11)
    sizes = np.rint(2.0 ** rng.uniform(10, 18, size=40)).astype(int)
    seeds = rng.integers(1, 1 << 40, size=40)
    arrivals = np.cumsum(rng.exponential(0.5, size=40))
    res = hz.Resident(keep_outputs=True)
    try:
        out = res.run_trace(list(zip(sizes.tolist(), seeds.tolist())), arrivals)
        for r, s, sd in zip(out, sizes, seeds):
            assert r["status"] == hz.JOB_DONE
            assert r["end_ms"] >= r["submit_ms"]
            check_job(res, r["id"], int(s), int(sd))
        res.drain(DRAIN_MS)
    finally:
        res.close()


def test_bad_requests_are_refused(hz):
    with pytest.raises(hz.HarmonizeError):
        hz.Resident(bits=41)
    res = hz.Resident(pool_bytes=12 * 4096, keep_outputs=True)
    try:
        with pytest.raises(hz.HarmonizeError):
            res.submit(0, 1)
        with pytest.raises(hz.HarmonizeError):
            res.submit(5000, 1)  # more items than the 4,096-item space holds
        j = hz._job("photon", 10, 1)
        i = ctypes.c_uint64()
        assert hz.lib().hz_resident_submit(res.handle, ctypes.byref(j), ctypes.byref(i)) != 0
        assert "Collatz" in hz.last_error()
    finally:
        res.close()


def test_jobs_wait_for_item_space_and_release_frees_it(hz):
    """An item space that holds one 3,000-item job at a time: three such jobs
    run one after another, in order, each correct; a released job's outputs
    are gone and its items serve the next job."""
    res = hz.Resident(pool_bytes=12 * 4096, keep_outputs=True)
    try:
        ids = [res.submit(3000, 40 + k) for k in range(3)]
        for k, i in enumerate(ids):
            deadline = time.monotonic() + 60.0
            while res.job(i)["status"] != hz.JOB_DONE:
                assert time.monotonic() < deadline, f"job {i} did not finish"
                res.poll()
            check_job(res, i, 3000, 40 + k)
            res.release(i)
            with pytest.raises(hz.HarmonizeError):
                res.output(i, 0)
        starts = [res.job(i)["start_ms"] for i in ids]
        ends = [res.job(i)["end_ms"] for i in ids]
        assert starts[1] >= ends[0] and starts[2] >= ends[1]
        res.drain(DRAIN_MS)
    finally:
        res.close()


# ------------------------------------------------------------------ photon jobs
# Events per photon bit-exact; tallies (FP64 atomics, summed in any order)
# within 1e-12 relative of the oracle's.

RTOL = 1e-12


def check_photon(res, i, size, seed):
    p = O.photon_params(photons=size, seed=seed)
    ref_t, ref_e, _ = O.photon_program(p, threads=0)
    assert np.array_equal(res.output(i, 0), ref_e), (i, size, seed)
    got = res.output(i, 1).reshape(-1, 1005).sum(axis=0)
    scale = np.maximum(np.abs(got), np.abs(ref_t))
    assert (np.abs(got - ref_t) <= RTOL * scale).all(), (i, size, seed)


@pytest.mark.parametrize("size", [1, 20000, (1 << 19) + 3, 1 << 22])
def test_photon_job_matches_oracle(hz, size):
    res = hz.Resident(kind="photon", keep_outputs=True)
    try:
        seed = 0xF0700 + size
        i = res.submit(size, seed)
        res.drain(DRAIN_MS)
        assert res.job(i)["status"] == hz.JOB_DONE
        check_photon(res, i, size, seed)
        assert res.info()["pool_bytes"] == 4 << 30  # 2^30 photons of 4 bytes
    finally:
        res.close()


def test_photon_jobs_reuse_table_entries(hz):
    """80 photon jobs at once: more than the 64 table entries, so entries
    (and their tallies) are reused while the launch runs; splits and roulette
    exercise the created/ended accounting of branching jobs."""
    rng = np.random.default_rng(3)
    sizes = np.rint(2.0 ** rng.uniform(4, 13, size=80)).astype(int)
    seeds = rng.integers(1, 1 << 40, size=80)
    res = hz.Resident(kind="photon", keep_outputs=True)
    try:
        ids = [res.submit(int(s), int(sd)) for s, sd in zip(sizes, seeds)]
        res.drain(DRAIN_MS)
        assert res.info()["joined_running"] > 0
        for i, s, sd in zip(ids, sizes, seeds):
            check_photon(res, i, int(s), int(sd))
    finally:
        res.close()


def test_photon_tally_lives_in_the_item_space(hz):
    """A photon job's allocation holds its tally (64 x 1005 doubles = 128,640
    words) before its event counts: a space of 100,000 words refuses even a
    10-photon job; without keep_outputs a finished job's tally is released
    with its event counts."""
    res = hz.Resident(kind="photon", pool_bytes=4 * 100000)
    try:
        with pytest.raises(hz.HarmonizeError, match="larger than the item space"):
            res.submit(10, 1)
    finally:
        res.close()
    res = hz.Resident(kind="photon")
    try:
        i = res.submit(3000, 9)
        res.drain(DRAIN_MS)
        assert res.job(i)["status"] == hz.JOB_DONE
        assert res.info()["pool_used"] == 0
        with pytest.raises(hz.HarmonizeError, match="released"):
            res.output(i, 1)
    finally:
        res.close()


def test_photon_smallest_first_is_refused(hz):
    with pytest.raises(hz.HarmonizeError, match="arrival order"):
        hz.Resident(kind="photon", policy=hz.RESIDENT_SMALLEST_FIRST)


def test_photon_open_loop_trace(hz):
    rng = np.random.default_rng(12)
    sizes = np.rint(2.0 ** rng.uniform(10, 17, size=30)).astype(int)
    seeds = rng.integers(1, 1 << 40, size=30)
    arrivals = np.cumsum(rng.exponential(0.3, size=30))
    res = hz.Resident(kind="photon", keep_outputs=True)
    try:
        out = res.run_trace(list(zip(sizes.tolist(), seeds.tolist())), arrivals)
        for r, s, sd in zip(out, sizes, seeds):
            assert r["status"] == hz.JOB_DONE
            check_photon(res, r["id"], int(s), int(sd))
        with pytest.raises(hz.HarmonizeError):
            res.submit(0, 1)
        j = hz._job("collatz", 10, 1)
        k = ctypes.c_uint64()
        assert hz.lib().hz_resident_submit(res.handle, ctypes.byref(j), ctypes.byref(k)) != 0
        res.drain(DRAIN_MS)
    finally:
        res.close()
