"""The serving loop on the GPU: every job type's output against the oracle,
each job's result identical alone and in a concurrent mix (photon tallies
within 1e-12 relative: their accumulation order differs), jobs of different
slots overlapping in time, and admission held back by pool memory."""

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

RTOL = 1e-12


@pytest.fixture(scope="module")
def hz():
    import harmonize_amd

    return harmonize_amd


def tally_sum(replicas):
    return replicas.reshape(-1, 1005).sum(axis=0)


def assert_close(a, b):
    scale = np.maximum(np.abs(a), np.abs(b))
    assert (np.abs(a - b) <= RTOL * scale).all()


def check_against_oracle(hz, srv, i, kind, size, seed):
    if kind == "collatz":
        starts = O.collatz_make_starts(seed, 32, size)
        assert np.array_equal(srv.output(i, 1), starts)
        want, promises = O.collatz_steps(starts)
        assert np.array_equal(srv.output(i, 0), want)
        assert srv.job(i)["promises"] == promises
    elif kind == "photon":
        p = O.photon_params(photons=size, seed=seed)
        ref_t, ref_e, ref_st = O.photon_program(p, threads=0)
        assert np.array_equal(srv.output(i, 0), ref_e)
        assert_close(tally_sum(srv.output(i, 1)), ref_t)
        assert srv.job(i)["promises"] == ref_st.promises
    else:
        r, c, e = O.bfs_csr(seed, np.array([size], np.uint32))
        assert np.array_equal(srv.output(i, 1), r)
        assert np.array_equal(srv.output(i, 2), c)
        assert np.array_equal(srv.output(i, 0), O.bfs_levels([size], e, r, c))


@pytest.mark.parametrize("kind,size", [("collatz", 1), ("collatz", 5000), ("collatz", (1 << 20) + 7),
                                       ("collatz", 1 << 22),
                                       ("photon", 1), ("photon", 20000), ("photon", (1 << 19) + 3),
                                       ("photon", 1 << 22),
                                       ("bfs", 1), ("bfs", 3000), ("bfs", 1 << 22)])
def test_job_matches_oracle(hz, kind, size):
    srv = hz.Server(pool_bytes=4 << 30, slots=2, keep_outputs=True)
    try:
        i = srv.submit(kind, size, 0xA11CE + size)
        srv.drain()
        assert srv.job(i)["status"] == hz.JOB_DONE
        check_against_oracle(hz, srv, i, kind, size, 0xA11CE + size)
    finally:
        srv.close()


def make_mix(n, seed, lo=10, hi=16):
    rng = np.random.default_rng(seed)
    kinds = rng.choice(["collatz", "photon", "bfs"], size=n)
    sizes = np.rint(2.0 ** rng.uniform(lo, hi, size=n)).astype(int)
    seeds = rng.integers(1, 1 << 40, size=n)
    return [(str(k), int(s), int(sd)) for k, s, sd in zip(kinds, sizes, seeds)]


def outputs(srv, i, kind):
    if kind == "photon":
        return [srv.output(i, 0), tally_sum(srv.output(i, 1))]
    return [srv.output(i, 0)]


def test_alone_equals_mix(hz):
    jobs = make_mix(30, 7)
    srv = hz.Server(pool_bytes=4 << 30, slots=4, keep_outputs=True)
    try:
        alone = []
        for kind, size, seed in jobs:
            i = srv.submit(kind, size, seed)
            srv.drain()
            alone.append(outputs(srv, i, kind))
            srv.release(i)
        # All at once (arrivals within 2 ms): every slot busy, jobs overlapping.
        arrivals = np.linspace(0.0, 2.0, len(jobs))
        res = srv.run_trace(jobs, arrivals)
        assert all(r["status"] == hz.JOB_DONE for r in res)
        for (kind, size, seed), r, a in zip(jobs, res, alone):
            got = outputs(srv, r["id"], kind)
            assert np.array_equal(got[0], a[0]), (kind, size, seed)
            if kind == "photon":
                assert_close(got[1], a[1])
        # Jobs on different slots ran at the same time.
        spans = sorted((r["start_ms"], r["end_ms"], r["slot"]) for r in res)
        overlaps = sum(1 for a, b in zip(spans, spans[1:]) if b[0] < a[1] and b[2] != a[2])
        assert overlaps > 0
        assert len({r["slot"] for r in res}) == 4
        check_against_oracle(hz, srv, res[0]["id"], *jobs[0])
    finally:
        srv.close()


def test_alone_equals_poisson_mix_full_size_range(hz):
    """The serving mix's own law: sizes log-uniform over [2^10, 2^22] items,
    Poisson arrivals (here at ~400 jobs/s, so jobs queue and overlap); every
    result identical to the same job run alone."""
    jobs = make_mix(24, 11, lo=10, hi=22)
    rng = np.random.default_rng(12)
    arrivals = np.cumsum(rng.exponential(2.5, size=len(jobs)))
    srv = hz.Server(pool_bytes=8 << 30, slots=3, keep_outputs=True)
    try:
        alone = []
        for kind, size, seed in jobs:
            i = srv.submit(kind, size, seed)
            srv.drain()
            alone.append(outputs(srv, i, kind))
            srv.release(i)
        res = srv.run_trace(jobs, arrivals)
        assert all(r["status"] == hz.JOB_DONE for r in res)
        for (kind, size, seed), r, a in zip(jobs, res, alone):
            got = outputs(srv, r["id"], kind)
            assert np.array_equal(got[0], a[0]), (kind, size, seed)
            if kind == "photon":
                assert_close(got[1], a[1])
        big = max(range(len(jobs)), key=lambda k: jobs[k][1])
        assert jobs[big][1] > 1 << 20
        check_against_oracle(hz, srv, res[big]["id"], *jobs[big])
    finally:
        srv.close()


def test_admission_waits_for_memory(hz):
    # A pool that holds one of these BFS jobs at a time: the second is admitted
    # only when the first has finished and released its state.
    size = 1 << 16
    srv = hz.Server(pool_bytes=0, slots=4)
    need = None
    try:
        i = srv.submit("bfs", size, 1)
        srv.drain()
        need = srv.job(i)["state_bytes"]
    finally:
        srv.close()
    srv = hz.Server(pool_bytes=int(need * 1.5), slots=4)
    try:
        res = srv.run_trace([("bfs", size, 1), ("bfs", size, 2), ("collatz", 100, 3)], [0.0, 0.0, 0.0])
        a, b, c = res
        assert b["start_ms"] >= a["end_ms"]
        assert c["start_ms"] >= a["end_ms"]  # arrival order is kept: it waits behind b
        assert srv.info()["pool_used"] == 0
    finally:
        srv.close()


def test_rejects_bad_jobs(hz):
    srv = hz.Server(pool_bytes=1 << 24, slots=1)
    try:
        with pytest.raises(hz.HarmonizeError):
            srv.submit("bfs", 1 << 23, 1)  # more vertices than a BFS job may have
        with pytest.raises(hz.HarmonizeError):
            srv.submit("collatz", 0, 1)
        with pytest.raises(hz.HarmonizeError):
            srv.submit("collatz", 1 << 22, 1)  # state larger than the pool
        i = srv.submit("collatz", 1000, 1)
        srv.drain()
        assert srv.job(i)["status"] == hz.JOB_DONE
    finally:
        srv.close()


def test_largest_first_policy(hz):
    # One slot, three jobs queued together: largest-first admits the largest
    # job first; arrival order admits them as they came.
    jobs = [("collatz", 2000, 1), ("collatz", 200000, 2), ("photon", 30000, 3)]
    order = {}
    for policy in (hz.SERVE_FIFO, hz.SERVE_LARGEST_FIRST):
        srv = hz.Server(pool_bytes=1 << 30, slots=1, policy=policy, keep_outputs=True)
        try:
            blocker = srv.submit("collatz", 50000, 9)  # occupies the slot while the others queue
            res = srv.run_trace(jobs, [0.0, 0.0, 0.0])
            order[policy] = [j for _, j in sorted((r["start_ms"], i) for i, r in enumerate(res))]
            want, _ = O.collatz_steps(O.collatz_make_starts(2, 32, 200000))
            assert np.array_equal(srv.output(res[1]["id"], 0), want)
            assert srv.job(blocker)["status"] == hz.JOB_DONE
        finally:
            srv.close()
    assert order[hz.SERVE_FIFO] == [0, 1, 2]
    assert order[hz.SERVE_LARGEST_FIRST] == [1, 2, 0]
