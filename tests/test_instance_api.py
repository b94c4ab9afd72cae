"""The generic instance entry points (hz_instance_create and friends): argument
checks on the CPU, and on the GPU a Collatz program driven through them (both
kinds) against the oracle, with a state round trip; the serving poll."""

import ctypes

import numpy as np
import pytest

import oracle_lib as O


def test_instance_arguments_are_checked():
    import harmonize_amd as hz

    st = hz.CollatzState()
    with pytest.raises(hz.HarmonizeError, match="no such specification"):
        hz.Instance("event", hz.SPEC_BARRIER, 1 << 10, st)  # the barrier program is async only
    with pytest.raises(hz.HarmonizeError, match="no such specification"):
        hz.Instance("async", 7, 1 << 10, st)
    with pytest.raises(hz.HarmonizeError, match="state must be"):
        hz.Instance("async", hz.SPEC_PHOTON, 1 << 10, st)  # a Collatz state for the photon program


class Iter(ctypes.Structure):
    _fields_ = [("value", ctypes.c_uint32), ("limit", ctypes.c_uint32)]


@pytest.mark.gpu
@pytest.mark.parametrize("kind,size,cycles", [("async", 1 << 12, 1 << 20), ("async", 1 << 12, 4), ("event", 1 << 18, 4)])
def test_collatz_through_instance(kind, size, cycles):
    import harmonize_amd as hz

    n = 20000
    host = O.collatz_make_starts(5, 40, n)
    starts = hz.DeviceArray.from_numpy(host)
    steps = hz.DeviceArray(n, np.uint32)
    it = hz.DeviceArray(2, np.uint32)
    it.upload(np.array([0, n], np.uint32))
    st = hz.CollatzState(starts.ptr, steps.ptr, it.ptr)
    inst = hz.Instance(kind, hz.SPEC_COLLATZ, size, st)
    try:
        back = inst.load_state()
        assert (back.starts, back.steps, back.iterator) == (st.starts, st.steps, st.iterator)
        inst.init(256)
        for _ in range(100000):
            inst.exec(256, cycles)
            if inst.complete():
                break
        else:
            raise AssertionError("the program did not complete")
        want, _ = O.collatz_steps(host)
        assert np.array_equal(steps.numpy(), want)
    finally:
        inst.close()


@pytest.mark.gpu
@pytest.mark.parametrize("kind,size,cycles", [("async", 1 << 12, 64), ("event", 1 << 18, 4)])
def test_two_instances_on_their_own_streams(kind, size, cycles):
    # Launches of two instances are enqueued on two streams without waiting,
    # alternating, until both report completion; each matches the oracle.
    import harmonize_amd as hz

    runs = []
    streams = [hz.stream_create(), hz.stream_create()]
    try:
        for seed, n in ((6, 30000), (7, 12345)):
            host = O.collatz_make_starts(seed, 40, n)
            starts = hz.DeviceArray.from_numpy(host)
            steps = hz.DeviceArray(n, np.uint32)
            it = hz.DeviceArray.from_numpy(np.array([0, n], np.uint32))
            st = hz.CollatzState(starts.ptr, steps.ptr, it.ptr)
            runs.append([host, starts, steps, it, hz.Instance(kind, hz.SPEC_COLLATZ, size, st), False])
        for r, s in zip(runs, streams):
            r[4].init_on(128, s)
        for _ in range(100000):
            for r, s in zip(runs, streams):
                if not r[5]:
                    r[4].exec_on(128, cycles, s)
            for r, s in zip(runs, streams):
                if not r[5]:
                    r[5] = r[4].complete_on(s)
            if all(r[5] for r in runs):
                break
        else:
            raise AssertionError("the programs did not complete")
        for host, _, steps, _, inst, _ in runs:
            want, _ = O.collatz_steps(host)
            assert np.array_equal(steps.numpy(), want)
            inst.close()
    finally:
        for s in streams:
            hz.stream_destroy(s)


@pytest.mark.gpu
def test_serve_poll_reports_each_job_once():
    import harmonize_amd as hz

    srv = hz.Server(pool_bytes=1 << 28, slots=2)
    try:
        ids = {srv.submit("collatz", 3000, 1), srv.submit("photon", 2000, 2), srv.submit("bfs", 1000, 3)}
        seen = []
        for _ in range(200000):
            seen += hz.serve_poll(srv)
            if len(seen) >= 3:
                break
        assert sorted(seen) == sorted(ids)
        assert hz.serve_poll(srv) == []
    finally:
        srv.close()
