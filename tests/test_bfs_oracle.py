"""The BFS oracle pinned on the CPU: the R-MAT generator against a numpy
restatement and its quadrant probabilities, the textbook BFS against brute
force on small graphs, and the BFS specification on the host executors
against the textbook BFS. No GPU."""

import numpy as np
import pytest

import oracle_lib as O

RMAT_STREAM = 0x524D4154
SIZE_STREAM = 0x53495A45
T_A, T_AB, T_ABC = 37356, 49807, 62259


def rmat_edge_restated(seed, g, e, vertices):
    """R-MAT edge from the published recursion: per level, a 16-bit uniform
    picks quadrant a/b/c/d (cumulative 0.57/0.76/0.95 of 2^16), appending a row
    and a column bit; Philox draws hold eight 16-bit uniforms, low half first."""
    scale = max(0, int(vertices - 1).bit_length())
    u = v = 0
    level = draw = 0
    while level < scale:
        words = O.philox([e, g, draw, RMAT_STREAM], [seed & 0xFFFFFFFF, seed >> 32])
        for j in range(8):
            if level == scale:
                break
            x = (words[j // 2] >> (16 * (j % 2))) & 0xFFFF
            q = 0 if x < T_A else 1 if x < T_AB else 2 if x < T_ABC else 3
            u = (u << 1) | (q >> 1)
            v = (v << 1) | (q & 1)
            level += 1
        draw += 1
    return (u * vertices) >> scale, (v * vertices) >> scale


@pytest.mark.parametrize("vertices", [1, 2, 3, 1000, 1 << 14, 3_000_001, 1 << 22])
def test_rmat_edge_matches_restatement(vertices):
    seed = 0x5EED
    for g, e in [(0, 0), (0, 1), (5, 77), (63, 12345), (17, (1 << 26) - 1)]:
        assert O.rmat_edge(seed, g, e, vertices) == rmat_edge_restated(seed, g, e, vertices)


def test_rmat_quadrant_probabilities():
    # A 2-vertex graph is one level of the recursion: its edge multiset gives
    # the quadrant frequencies directly.
    n = 1 << 20
    row_ptr, col, _ = O.bfs_csr(11, np.array([2], np.uint32), edge_factor=n // 2)
    src = np.repeat(np.arange(2), np.diff(row_ptr[:3]))
    counts = np.zeros(4)
    np.add.at(counts, 2 * src + col, 1)
    freq = counts / n
    sd = np.sqrt(0.25 / n)
    assert np.allclose(freq, [0.57, 0.19, 0.19, 0.05], atol=6 * sd)


def test_rmat_skew_and_range():
    v = np.array([1 << 12], np.uint32)
    row_ptr, col, edges = O.bfs_csr(3, v)
    deg = np.diff(row_ptr)
    assert row_ptr[0] == 0 and row_ptr[-1] == edges[0] == 16 * v[0]
    assert (deg >= 0).all() and (col >= 0).all() and (col < v[0]).all()
    assert deg[0] == deg.max()  # vertex 0 is the R-MAT hub


def test_batch_sizes_match_restatement():
    import harmonize_amd as hz

    seed, graphs = 0x5EED, 64
    got = hz.bfs_batch_sizes(seed, graphs, 14.0, 22.0)
    want = []
    for g in range(graphs):
        bits = O.philox([g, 0, 0, SIZE_STREAM], [seed & 0xFFFFFFFF, seed >> 32])[0]
        want.append(int(np.rint(2.0 ** (14.0 + 8.0 * (bits / 2.0**32)))))
    assert got.tolist() == want
    assert got.min() >= 1 << 14 and got.max() <= 1 << 22


def brute_levels(V, edge_list, source):
    """Bellman-Ford style relaxation to a fixed point on unit weights."""
    INF = O.BFS_UNREACHED
    lev = np.full(V, INF, dtype=np.int64)
    if source < V:
        lev[source] = 0
    for _ in range(V + 1):
        changed = False
        for a, b in edge_list:
            if lev[a] != INF and lev[a] + 1 < lev[b]:
                lev[b] = lev[a] + 1
                changed = True
        if not changed:
            break
    return lev.astype(np.int32)


def to_csr(V, edge_list):
    edge_list = sorted(edge_list, key=lambda t: t[0])  # stable: keeps insertion order per source
    row = np.zeros(V + 1, np.int32)
    for a, _ in edge_list:
        row[a + 1] += 1
    row = np.cumsum(row).astype(np.int32)
    col = np.array([b for _, b in edge_list], np.int32)
    return row, col


def batch_of(graphs):
    """[(V, edges, source)] -> concatenated batch arrays."""
    vertices, edges, rows, cols, sources = [], [], [], [], []
    for V, el, s in graphs:
        r, c = to_csr(V, el)
        vertices.append(V)
        edges.append(len(el))
        rows.append(r)
        cols.append(c)
        sources.append(s)
    return (np.array(vertices, np.uint32), np.array(edges, np.uint64), np.concatenate(rows),
            np.concatenate(cols) if sum(edges) else np.zeros(0, np.int32), np.array(sources, np.uint32))


def small_graphs(seed):
    rng = np.random.default_rng(seed)
    out = [
        (1, [], 0),                                   # single vertex, no edges
        (5, [], 2),                                   # no edges, source in the middle
        (4, [(0, 0), (0, 0), (1, 2)], 0),             # self-loops and duplicates, 1..3 unreachable
        (6, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (0, 5)], 0),  # shortcut beats the chain
        (6, [(5, 4), (4, 3), (3, 2), (2, 1), (1, 0)], 5),          # source is the last vertex
    ]
    for _ in range(20):
        V = int(rng.integers(2, 40))
        E = int(rng.integers(0, 4 * V))
        el = [(int(a), int(b)) for a, b in rng.integers(0, V, size=(E, 2))]
        out.append((V, el, int(rng.integers(0, V))))
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_textbook_bfs_matches_brute_force(seed):
    graphs = small_graphs(seed)
    v, e, r, c, s = batch_of(graphs)
    got = O.bfs_levels(v, e, r, c, s)
    want = np.concatenate([brute_levels(V, el, src) for V, el, src in graphs])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("threads", [1, 0])
@pytest.mark.parametrize("split", [1, 2, 64])
def test_spec_on_host_matches_textbook_small(threads, split):
    v, e, r, c, s = batch_of(small_graphs(7))
    want = O.bfs_levels(v, e, r, c, s)
    got, st = O.bfs_program(v, e, r, c, s, split=split, threads=threads)
    assert np.array_equal(got, want)
    assert st.promises >= int((want < O.BFS_UNREACHED).sum())  # at least one visit per reached vertex


@pytest.mark.parametrize("threads", [1, 0])
def test_spec_on_host_matches_textbook_rmat(threads):
    v = np.array([1, 2, 100, 1 << 10, 3000, 1 << 13], np.uint32)
    r, c, e = O.bfs_csr(0x5EED, v)
    want = O.bfs_levels(v, e, r, c)
    got, _ = O.bfs_program(v, e, r, c, threads=threads)
    assert np.array_equal(got, want)
    assert (want[:1] == 0).all()
