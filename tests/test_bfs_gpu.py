"""Batched BFS on the GPU: CSR built on the device bit-exact against the
oracle's host build, and int32 levels bit-exact against the textbook BFS for
both program types, every launch shape the runtime offers (workgroup sizes,
queues, coalescing, split threshold, relaunches, graph capture), hand-made
edge-case graphs, and the 64-graph R-MAT batch of 2^14..2^22 vertices."""

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

SEED = 0x5EED


@pytest.fixture(scope="module")
def hz():
    import harmonize_amd

    return harmonize_amd


@pytest.fixture(scope="module")
def small_batch(hz):
    sizes = hz.bfs_batch_sizes(SEED, 16, 6.0, 14.0)
    b = hz.BfsBatch.rmat(SEED, sizes)
    row_ptr, col = b.csr()
    want = O.bfs_levels(b.vertices, b.edges, row_ptr, col)
    yield b, want
    b.close()


def test_csr_matches_host_build(hz, small_batch):
    b, _ = small_batch
    r_host, c_host, e_host = O.bfs_csr(SEED, b.vertices)
    r_dev, c_dev = b.csr()
    assert np.array_equal(b.edges, e_host)
    assert np.array_equal(r_dev, r_host)
    assert np.array_equal(c_dev, c_host)


@pytest.mark.parametrize(
    "kind,cfg",
    [
        ("async", {}),
        ("async", {"block_groups": 2}),
        ("async", {"block_groups": 4}),
        ("async", {"stash": 8}),
        ("async", {"stash": 8, "block_groups": 4, "flags": 1}),
        ("async", {"queues": 1, "coalesce": 16}),
        ("async", {"spec_param": 1}),
        ("async", {"spec_param": 1024, "queues": 16}),
        ("async", {"groups": 64, "arena": 1 << 14}),
        ("async", {"flags": 1}),
        ("event", {}),
        ("event", {"spec_param": 4, "cycles": 1}),
    ],
)
def test_levels_match_textbook(hz, small_batch, kind, cfg):
    b, want = small_batch
    st = b.run(kind, max_launches=1 << 16, **cfg)
    got = b.levels()
    assert np.array_equal(got, want)
    assert st["promises"] >= int((want < O.BFS_UNREACHED).sum())


def test_relaunch_resumes(hz, small_batch):
    b, want = small_batch
    st = b.run("async", groups=32, cycles=8, max_launches=1 << 20, flags=hz.RUN_AUDIT)
    assert st["launches"] > 1 and st["audits"] == st["launches"]
    assert np.array_equal(b.levels(), want)


def test_graph_replays(hz, small_batch):
    """A captured pass replayed several times, with other work in between;
    every replay must give the same levels."""
    b, want = small_batch
    for i in range(4):
        st = b.run("async", flags=hz.RUN_GRAPH, max_launches=1 << 16)
        assert np.array_equal(b.levels(), want), f"replay {i}"
        b.run("event")  # eager work between replays
        assert np.array_equal(b.levels(), want)
    b.release()


def test_edge_case_graphs(hz):
    from test_bfs_oracle import batch_of, small_graphs

    graphs = small_graphs(5)
    v, e, r, c, s = batch_of(graphs)
    want = O.bfs_levels(v, e, r, c, s)
    b = hz.BfsBatch.from_csr(v, e, r, c, s)
    for kind in ("async", "event"):
        b.run(kind, max_launches=1 << 16)
        assert np.array_equal(b.levels(), want), kind
    b.close()


def test_from_csr_rejects_bad_input(hz):
    with pytest.raises(hz.HarmonizeError):
        hz.BfsBatch.from_csr([3], [2], [0, 1, 2, 2], [0, 7])  # column out of range
    with pytest.raises(hz.HarmonizeError):
        hz.BfsBatch.from_csr([3], [2], [0, 2, 1, 2], [0, 1])  # row_ptr decreases


@pytest.fixture(scope="module")
def full_batch(hz):
    sizes = hz.bfs_batch_sizes(SEED, 64, 14.0, 22.0)
    b = hz.BfsBatch.rmat(SEED, sizes)
    row_ptr, col = b.csr()
    want = O.bfs_levels(b.vertices, b.edges, row_ptr, col)
    yield b, sizes, row_ptr, col, want
    b.close()


def test_rmat_batch_64_graphs(hz, full_batch):
    """The 64-graph batch, 2^14..2^22 vertices, 16 edges per vertex: levels of
    both program types bit-exact against the textbook BFS on the same CSR; the
    CSR of the eight smallest graphs bit-exact against the host build."""
    b, sizes, row_ptr, col, want = full_batch
    for kind in ("async", "event"):
        b.run(kind, max_launches=1 << 16)
        assert np.array_equal(b.levels(), want), kind
    for g in np.argsort(sizes, kind="stable")[:8]:
        r_host, c_host, _ = O.bfs_csr(SEED, sizes[g:g + 1], first_graph=int(g))
        r_dev = row_ptr[b.row_base[g]: b.row_base[g] + sizes[g] + 1]
        c_dev = col[b.edge_base[g]: b.edge_base[g] + b.edges[g]]
        assert np.array_equal(r_dev, r_host) and np.array_equal(c_dev, c_host), f"graph {g}"


@pytest.mark.parametrize(
    "cfg",
    [
        # a configuration the benchmark's tuner chose for this batch
        {"groups": 2048, "arena": 466741, "queues": 16, "coalesce": 32, "block_groups": 2, "spec_param": 128,
         "stash": 8},
        {"groups": 2048, "arena": 466741, "queues": 16, "coalesce": 32, "block_groups": 4, "spec_param": 64,
         "stash": 16},
    ],
)
def test_rmat_batch_64_graphs_captured_at_tuned_shapes(hz, full_batch, cfg):
    """The 64-graph batch through the captured hipGraph pass at the shapes the
    benchmark times, replayed twice with an eager pass in between."""
    b, _, _, _, want = full_batch
    for replay in range(2):
        b.run("async", flags=hz.RUN_GRAPH, max_launches=1 << 16, **cfg)
        assert np.array_equal(b.levels(), want), f"replay {replay}"
        b.run("event")
        assert np.array_equal(b.levels(), want)
    b.release()


def test_rmat_batch_64_graphs_at_the_tuners_choice(hz, full_batch):
    """The tuner's own pick on the 64-graph batch (one round of single trials),
    captured and replayed, against the textbook BFS."""
    b, _, _, _, want = full_batch
    res = hz.tune_bfs(b, rounds=1, reps=1)
    best = {k: v for k, v in res["best"].items() if v}
    assert best.get("flags", 0) & hz.RUN_GRAPH
    best.setdefault("max_launches", 1 << 16)
    for replay in range(2):
        b.run("async", **best)
        assert np.array_equal(b.levels(), want), f"replay {replay} of {best}"
    b.release()


def test_tuner_picks_a_working_configuration(hz, small_batch):
    b, want = small_batch
    res = hz.tune_bfs(b, rounds=1, reps=1)
    assert res["trials"] == len(res["log"]) and res["trials"] > 10
    assert res["best_ms"] > 0 and (res["start_ms"] < 0 or res["best_ms"] <= res["start_ms"])
    best = {k: v for k, v in res["best"].items() if v}
    st = b.run("async", **best)
    assert np.array_equal(b.levels(), want)
    assert st["launches"] >= 1
    b.release()


def test_tuner_cache_returns_the_stored_choice(hz, small_batch, tmp_path):
    """A second tuning of the same batch class on the same device comes from the
    cache: no trial, the same configuration, which still computes the levels."""
    b, want = small_batch
    hz.tune_cache_set_path(str(tmp_path / "tune.txt"))
    try:
        first = hz.tune_bfs(b, rounds=1, reps=1)
        second = hz.tune_bfs(b, rounds=1, reps=1)
        assert not first["cached"] and first["trials"] > 10
        assert second["cached"] and second["trials"] == 0 and second["best"] == first["best"]
        best = {k: v for k, v in second["best"].items() if v}
        best.setdefault("max_launches", 1 << 16)
        b.run("async", **best)
        assert np.array_equal(b.levels(), want)
    finally:
        hz.tune_cache_set_path(None)
        b.release()
