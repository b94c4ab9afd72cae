"""The depth-level tree specification on the oracle's host executors, pinned
by closed forms: root r counts fanout^height leaves, the run evaluates
roots * (fanout^(height+1) - 1) / (fanout - 1) promises, and every node's
reported depth is its height clamped to the program's levels (the spec
counts mismatches). A planted mistake in the level arithmetic (a depth delta
that is dropped, or not clamped) shows up as violations."""

import pytest

import oracle_lib as O


def promises_of(roots, fanout, height):
    if fanout == 1:
        return roots * (height + 1)
    return roots * (fanout ** (height + 1) - 1) // (fanout - 1)


@pytest.mark.parametrize("order", [O.TREE_FLAT, O.TREE_DEEPEST, O.TREE_SHALLOWEST])
@pytest.mark.parametrize("threads", [1, 0])
@pytest.mark.parametrize("roots,fanout,height", [(1, 1, 0), (3, 2, 5), (100, 3, 6), (17, 2, 11), (5, 1, 20)])
def test_tree_closed_forms(order, threads, roots, fanout, height):
    leaves, violations, st = O.tree_program(order, roots, fanout, height, threads=threads)
    assert (leaves == fanout ** height).all()
    assert violations == 0
    assert st.promises == promises_of(roots, fanout, height)


def test_depth_levels_clamp_at_the_stack():
    # Heights 8..11 run at the deepest level (7): the spec's own check asks for
    # min(height, 7), so a level that kept growing past the stack would count
    # violations; a flat program reports level 0 everywhere.
    for order in (O.TREE_DEEPEST, O.TREE_SHALLOWEST, O.TREE_FLAT):
        _, violations, _ = O.tree_program(order, 4, 2, 11)
        assert violations == 0


def test_zero_roots():
    leaves, violations, st = O.tree_program(O.TREE_DEEPEST, 0, 4, 3)
    assert leaves.size == 0 and violations == 0 and st.promises == 0
