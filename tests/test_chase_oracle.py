"""The pointer-chasing specification (LazyLoad arguments) on the oracle's host
executors, pinned by the closed form: each list's sum and length computed
directly from the permutation that built it. Both forms of the argument agree:
a LazyLoad<node> parameter (read when the promise is evaluated) and a plain
node parameter (read when the promise is built)."""

import numpy as np
import pytest

import oracle_lib as O


@pytest.mark.parametrize("lazy", [0, 1])
@pytest.mark.parametrize("threads", [1, 0])
@pytest.mark.parametrize("nodes,lists", [(0, 1), (1, 1), (10, 3), (1000, 64), (50000, 777)])
def test_chase_closed_form(lazy, threads, nodes, lists):
    heads, table, want_sums, want_lengths = O.chase_lists(nodes, lists, seed=nodes + lists)
    sums, lengths, st = O.chase_program(lazy, heads, table, threads=threads)
    assert np.array_equal(sums, want_sums)
    assert np.array_equal(lengths, want_lengths)
    assert st.promises == nodes


def test_closed_form_builder_itself():
    # Walking the built table by hand reproduces the builder's sums and lengths.
    heads, table, sums, lengths = O.chase_lists(300, 9, seed=5)
    seen = set()
    for l, h in enumerate(heads):
        total, n, i = 0, 0, int(h)
        while i != O.CHASE_END:
            assert i not in seen
            seen.add(i)
            total += int(table[i, 1])
            n += 1
            i = int(table[i, 0])
        assert total == int(sums[l]) and n == int(lengths[l])
    assert len(seen) == 300
