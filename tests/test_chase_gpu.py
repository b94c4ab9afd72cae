"""Asynchronous argument loads on the GPU (hz_chase_run): list sums and
lengths bit-exact against the closed form and the host executors, for the
LazyLoad parameter (nodes loaded ahead of the link that consumes them) and
the plain parameter (nodes read where the promise is built), under both
program types, with ragged and empty lists and across relaunches."""

import numpy as np
import pytest

import harmonize_amd as hz
import oracle_lib as O

pytestmark = pytest.mark.gpu


def upload(heads, table):
    return hz.DeviceArray.from_numpy(heads), hz.DeviceArray.from_numpy(table.reshape(-1))


@pytest.mark.parametrize("kind", ["async", "event"])
@pytest.mark.parametrize("lazy", [True, False])
@pytest.mark.parametrize("nodes,lists", [(1, 1), (100, 7), (5000, 64), (1 << 18, 4096), (1 << 20, 1 << 14)])
def test_chase_matches_closed_form(kind, lazy, nodes, lists):
    heads, table, want_sums, want_lengths = O.chase_lists(nodes, lists, seed=3 * nodes + lists)
    dh, dn = upload(heads, table)
    sums, lengths, st = hz.chase_run(kind, lazy, dh, dn, max_launches=1 << 20)
    assert np.array_equal(sums, want_sums)
    assert np.array_equal(lengths, want_lengths)
    assert st["promises"] == nodes


def test_chase_matches_host_executor():
    heads, table, _, _ = O.chase_lists(20000, 300, seed=9)
    ref_sums, ref_lengths, ref_st = O.chase_program(1, heads, table, threads=0)
    dh, dn = upload(heads, table)
    sums, lengths, st = hz.chase_run("async", True, dh, dn)
    assert np.array_equal(sums, ref_sums) and np.array_equal(lengths, ref_lengths)
    assert st["promises"] == ref_st.promises


@pytest.mark.parametrize("lazy", [True, False])
def test_chase_relaunches_with_audit(lazy):
    # Short launches spill prefetched and unprefetched links alike.
    heads, table, want_sums, want_lengths = O.chase_lists(1 << 16, 2000, seed=21)
    dh, dn = upload(heads, table)
    sums, lengths, st = hz.chase_run("async", lazy, dh, dn, groups=64, cycles=16, max_launches=1 << 20,
                                     flags=hz.RUN_AUDIT)
    assert st["launches"] > 1 and st["audits"] == st["launches"]
    assert np.array_equal(sums, want_sums) and np.array_equal(lengths, want_lengths)


@pytest.mark.parametrize("nodes,lists", [(1 << 24, 1 << 18), (1 << 22, 1 << 10)])
def test_chase_lazy_against_eager_timing(nodes, lists):
    # Every step is a dependent load from a random place in HBM: many short
    # lists (parallel) and few long ones (latency-bound). Reports both forms
    # (results must agree).
    heads, table, want_sums, want_lengths = O.chase_lists(nodes, lists, seed=77)
    dh, dn = upload(heads, table)
    ms = {}
    for lazy in (False, True):
        best = None
        for _ in range(3):
            sums, lengths, st = hz.chase_run("async", lazy, dh, dn)
            assert np.array_equal(sums, want_sums) and np.array_equal(lengths, want_lengths)
            best = st["elapsed_ms"] if best is None else min(best, st["elapsed_ms"])
        ms[lazy] = best
    ev = hz.chase_run("event", True, dh, dn, max_launches=1 << 20)[2]["elapsed_ms"]
    print(f"chase {nodes} nodes in {lists} lists: eager {ms[False]:.2f} ms, lazy {ms[True]:.2f} ms, "
          f"event {ev:.2f} ms", flush=True)
