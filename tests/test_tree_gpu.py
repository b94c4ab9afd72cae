"""Depth levels on the GPU (the tree of specs/tree.h through hz_tree_run):
for every order, each root counts fanout^height leaves, the promise count is
the closed form, every node reports its own depth level, and the queues and
arena pass the host audit across relaunches; the host executors agree. The
deepest-first order, the reference's, needs fewer arena links than flat or
shallowest-first order on the same tree."""

import numpy as np
import pytest

import harmonize_amd as hz
import oracle_lib as O

pytestmark = pytest.mark.gpu

ORDERS = [hz.TREE_FLAT, hz.TREE_DEEPEST, hz.TREE_SHALLOWEST]


def promises_of(roots, fanout, height):
    return roots * (fanout ** (height + 1) - 1) // (fanout - 1) if fanout > 1 else roots * (height + 1)


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("roots,fanout,height", [(1, 2, 0), (1, 3, 9), (1000, 2, 7), (4096, 4, 6)])
def test_tree_matches_closed_form(order, roots, fanout, height):
    leaves, violations, st = hz.tree_run("async", order, roots, fanout, height, max_launches=100000)
    assert (leaves == fanout ** height).all()
    assert violations == 0
    assert st["promises"] == promises_of(roots, fanout, height)
    ref_leaves, ref_violations, ref_st = O.tree_program(order, roots, fanout, height, threads=0)
    assert np.array_equal(leaves, ref_leaves) and ref_violations == 0 and ref_st.promises == st["promises"]


@pytest.mark.parametrize("order", ORDERS)
def test_tree_levels_past_the_stack_and_relaunches(order):
    # Heights beyond the 8 levels run at the deepest level; short launches
    # park every level's links in the arena between launches (64 groups x 4
    # cycles x at most 8 links of 64 promises is half the run's 262,080).
    roots, fanout, height = 64, 2, 11
    leaves, violations, st = hz.tree_run("async", order, roots, fanout, height, groups=64, cycles=4,
                                         max_launches=100000, flags=hz.RUN_AUDIT)
    assert st["launches"] > 1 and st["audits"] == st["launches"] and st["spills"] > 0
    assert (leaves == fanout ** height).all() and violations == 0
    assert st["promises"] == promises_of(roots, fanout, height)


def test_tree_under_event_program():
    leaves, violations, st = hz.tree_run("event", hz.TREE_FLAT, 500, 3, 6, max_launches=1 << 16)
    assert (leaves == 3 ** 6).all() and violations == 0
    assert st["promises"] == promises_of(500, 3, 6)


def test_deepest_first_bounds_the_arena():
    # 2^14 roots of 4-ary trees of height 7 (358 M promises). Breadth-first
    # expansion parks most of a level in the arena (shallowest first needs
    # more than 2^20 links of 512 B here), deepest first finishes subtrees.
    roots, fanout, height = 1 << 14, 4, 7
    used = {}
    for order in ORDERS:
        arena = 1 << 22 if order == hz.TREE_SHALLOWEST else 1 << 20
        leaves, violations, st = hz.tree_run("async", order, roots, fanout, height, arena=arena,
                                             max_launches=100000)
        assert (leaves == fanout ** height).all() and violations == 0
        used[order] = st["arena_links"]
        print(f"order {order}: arena links {st['arena_links']}, spills {st['spills']}, "
              f"{st['elapsed_ms']:.2f} ms", flush=True)
    assert used[hz.TREE_DEEPEST] < used[hz.TREE_FLAT]
    assert used[hz.TREE_DEEPEST] < used[hz.TREE_SHALLOWEST]
