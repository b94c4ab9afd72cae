"""ctypes handle on oracle/liboracle.so, the CPU oracle (test infrastructure only)."""

import ctypes
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_PATH = ROOT / "oracle" / "liboracle.so"

_VP = ctypes.c_void_p


class OracleStats(ctypes.Structure):
    _fields_ = [
        ("promises", ctypes.c_uint64),
        ("make_work_calls", ctypes.c_uint64),
        ("peak_queue", ctypes.c_uint64),
        ("elapsed_ms", ctypes.c_double),
        ("threads", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


_SIGS = {
    "oracle_collatz_steps_one": (ctypes.c_uint32, [ctypes.c_uint64]),
    "oracle_collatz_steps": (ctypes.c_uint64, [_VP, ctypes.c_uint32, _VP, ctypes.c_int]),
    "oracle_collatz_program": (ctypes.c_int, [_VP, ctypes.c_uint32, _VP, ctypes.c_int, ctypes.POINTER(OracleStats)]),
    "oracle_philox4x32_10": (None, [_VP, _VP, _VP]),
    "oracle_collatz_make_starts": (None, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _VP]),
    "oracle_max_threads": (ctypes.c_int, []),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not ORACLE_PATH.exists():
            raise RuntimeError(f"{ORACLE_PATH} is not built; run `python -m harmonize_amd.build --oracle`")
        h = ctypes.CDLL(str(ORACLE_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def collatz_steps(starts, count=None, threads=0):
    """Closed-form step counts; returns (steps, promises)."""
    n = len(starts) if starts is not None else int(count)
    if starts is not None:
        starts = np.ascontiguousarray(starts, dtype=np.uint64)
    out = np.zeros(n, dtype=np.uint32)
    promises = lib().oracle_collatz_steps(_ptr(starts), n, out.ctypes.data, threads)
    return out, int(promises)


def collatz_program(starts, count=None, threads=1):
    """Runs the Collatz specification on the host executor; returns (steps, stats)."""
    n = len(starts) if starts is not None else int(count)
    if starts is not None:
        starts = np.ascontiguousarray(starts, dtype=np.uint64)
    out = np.zeros(n, dtype=np.uint32)
    st = OracleStats()
    lib().oracle_collatz_program(_ptr(starts), n, out.ctypes.data, threads, ctypes.byref(st))
    return out, st


def collatz_make_starts(seed, bits, count):
    out = np.zeros(count, dtype=np.uint64)
    lib().oracle_collatz_make_starts(seed, bits, count, out.ctypes.data)
    return out


def philox(counter, key):
    c = np.asarray(counter, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().oracle_philox4x32_10(c.ctypes.data, k.ctypes.data, o.ctypes.data)
    return [int(x) for x in o]


class PhotonParams(ctypes.Structure):
    _fields_ = [
        ("photons", ctypes.c_uint32), ("bins", ctypes.c_uint32), ("seed", ctypes.c_uint64),
        ("thickness", ctypes.c_float), ("sigma_t", ctypes.c_float), ("albedo", ctypes.c_float),
        ("split_above", ctypes.c_float), ("roulette_below", ctypes.c_float), ("region_width", ctypes.c_float),
        ("regions", ctypes.c_uint32), ("max_split", ctypes.c_uint32), ("max_events", ctypes.c_uint32),
        ("replicas", ctypes.c_uint32),
    ]


PHOTON_DEFAULTS = dict(photons=10_000_000, bins=1000, seed=0x5EED, thickness=20.0, sigma_t=1.0, albedo=0.9,
                       split_above=2.0, roulette_below=0.1, region_width=2.0, regions=10, max_split=8,
                       max_events=65535, replicas=256)

_SIGS.update({
    "oracle_photon_program": (ctypes.c_int, [ctypes.POINTER(PhotonParams), ctypes.c_uint32, ctypes.c_uint32, _VP,
                                             _VP, ctypes.c_int, ctypes.POINTER(OracleStats)]),
    "oracle_det_log": (ctypes.c_double, [ctypes.c_double]),
    "oracle_det_sqrt": (ctypes.c_double, [ctypes.c_double]),
})


_SIGS["oracle_barrier_program"] = (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _VP, _VP, _VP,
                                                   ctypes.c_int, ctypes.POINTER(OracleStats)])


def barrier_program(dep, non_dep, barriers, threads=1):
    """Runs the barrier specification on the host executors; returns (counter, hits, violations, stats)."""
    counter = np.zeros(1, dtype=np.uint32)
    violations = np.zeros(1, dtype=np.uint32)
    hits = np.zeros(max(1, dep), dtype=np.uint32)
    st = OracleStats()
    lib().oracle_barrier_program(dep, non_dep, barriers, counter.ctypes.data, hits.ctypes.data,
                                 violations.ctypes.data, threads, ctypes.byref(st))
    return int(counter[0]), hits[:dep], int(violations[0]), st


_SIGS["oracle_tree_program"] = (ctypes.c_int, [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _VP,
                                                _VP, ctypes.c_int, ctypes.POINTER(OracleStats)])

TREE_FLAT, TREE_DEEPEST, TREE_SHALLOWEST = 0, 1, 2
TREE_LEVELS = 8


def tree_program(order, roots, fanout, height, threads=1):
    """Runs the depth-level tree specification on the host executors; returns (leaves, violations, stats)."""
    leaves = np.zeros(max(1, roots), dtype=np.uint32)
    violations = np.zeros(1, dtype=np.uint32)
    st = OracleStats()
    lib().oracle_tree_program(order, roots, fanout, height, leaves.ctypes.data, violations.ctypes.data, threads,
                              ctypes.byref(st))
    return leaves[:roots], int(violations[0]), st


_SIGS["oracle_chase_program"] = (ctypes.c_int, [ctypes.c_int, ctypes.c_uint32, _VP, _VP, _VP, _VP, ctypes.c_int,
                                                 ctypes.POINTER(OracleStats)])
CHASE_END = 0xFFFFFFFF


def chase_lists(nodes, lists, seed):
    """Random linked lists: a permutation of `nodes` node ids cut into `lists`
    ragged runs (some empty); returns (heads[lists] u32, node table
    [nodes][next, value] u32, sums u64, lengths u32). The sums and lengths are
    the closed form: each run's values summed in order."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(nodes).astype(np.uint32)
    cuts = np.sort(rng.integers(0, nodes + 1, size=max(0, lists - 1)))
    bounds = np.concatenate([[0], cuts, [nodes]]).astype(np.int64)
    values = rng.integers(0, 1 << 32, size=nodes, dtype=np.uint64).astype(np.uint32)
    table = np.zeros((nodes, 2), dtype=np.uint32)
    table[:, 1] = values
    heads = np.full(lists, CHASE_END, dtype=np.uint32)
    sums = np.zeros(lists, dtype=np.uint64)
    lengths = (bounds[1:] - bounds[:-1]).astype(np.uint32)
    for l in range(lists):
        a, b = bounds[l], bounds[l + 1]
        if b > a:
            run = perm[a:b]
            heads[l] = run[0]
            table[run[:-1], 0] = run[1:]
            table[run[-1], 0] = CHASE_END
            sums[l] = values[run].astype(np.uint64).sum()
    return heads, table, sums, lengths


def chase_program(lazy, heads, table, threads=1):
    """Runs the pointer-chasing specification on the host executors; returns (sums, lengths, stats)."""
    lists = len(heads)
    heads = np.ascontiguousarray(heads, dtype=np.uint32)
    table = np.ascontiguousarray(table, dtype=np.uint32)
    sums = np.zeros(max(1, lists), dtype=np.uint64)
    lengths = np.zeros(max(1, lists), dtype=np.uint32)
    st = OracleStats()
    lib().oracle_chase_program(int(lazy), lists, heads.ctypes.data, table.ctypes.data, sums.ctypes.data,
                               lengths.ctypes.data, threads, ctypes.byref(st))
    return sums[:lists], lengths[:lists], st


_SIGS["oracle_futures_program"] = (ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.c_uint32, _VP, _VP, _VP, ctypes.c_int,
                                                   ctypes.POINTER(OracleStats)])

FUTURES_EARLY = (1 << 64) - 1


def futures_program(starts, readers, threads=1):
    """Runs the futures specification on the host executors over `starts` (a
    power-of-two count); returns (total, echo[readers], forwarded[readers], stats)."""
    starts = np.ascontiguousarray(starts, dtype=np.uint64)
    echo = np.zeros(max(1, readers), dtype=np.uint64)
    forwarded = np.zeros(max(1, readers), dtype=np.uint64)
    total = np.zeros(1, dtype=np.uint64)
    st = OracleStats()
    rc = lib().oracle_futures_program(starts.ctypes.data, len(starts), readers, echo.ctypes.data,
                                      forwarded.ctypes.data, total.ctypes.data, threads, ctypes.byref(st))
    if rc != 0:
        raise ValueError("futures_program: the leaf count must be a power of two")
    return int(total[0]), echo[:readers], forwarded[:readers], st


def photon_params(**overrides):
    d = dict(PHOTON_DEFAULTS)
    d.update(overrides)
    return PhotonParams(**d)


def photon_program(params, first=0, count=None, threads=1):
    """Runs photons [first, first+count) on the host executors; returns (tally, events[photons], stats)."""
    count = params.photons - first if count is None else count
    tally = np.zeros(params.bins + 5)
    events = np.zeros(params.photons, dtype=np.uint32)
    st = OracleStats()
    lib().oracle_photon_program(ctypes.byref(params), first, count, tally.ctypes.data, events.ctypes.data, threads,
                                ctypes.byref(st))
    return tally, events, st


_SIGS.update({
    "oracle_rmat_edge": (None, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _VP, _VP]),
    "oracle_bfs_csr": (None, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _VP, ctypes.c_uint32, _VP, _VP]),
    "oracle_bfs_levels": (None, [ctypes.c_uint32, _VP, _VP, _VP, _VP, _VP, _VP, ctypes.c_int]),
    "oracle_bfs_program": (ctypes.c_int, [ctypes.c_uint32, _VP, _VP, _VP, _VP, _VP, _VP, ctypes.c_uint32, ctypes.c_int,
                                          ctypes.POINTER(OracleStats)]),
})

BFS_UNREACHED = 0x7FFFFFFF


def rmat_edge(seed, g, e, vertices):
    s = ctypes.c_uint32()
    d = ctypes.c_uint32()
    lib().oracle_rmat_edge(seed, g, e, vertices, ctypes.byref(s), ctypes.byref(d))
    return s.value, d.value


def bfs_csr(seed, vertices, edge_factor=16, first_graph=0):
    """Host-built CSR of R-MAT graphs first_graph, first_graph + 1, ... of a
    batch: (row_ptr, col, edges per graph)."""
    v = np.ascontiguousarray(vertices, dtype=np.uint32)
    edges = v.astype(np.uint64) * edge_factor
    row_ptr = np.zeros(int(v.sum()) + len(v), dtype=np.int32)
    col = np.zeros(int(edges.sum()), dtype=np.int32)
    lib().oracle_bfs_csr(seed, first_graph, len(v), v.ctypes.data, edge_factor, row_ptr.ctypes.data,
                         col.ctypes.data)
    return row_ptr, col, edges


def _bfs_args(vertices, edges, row_ptr, col, sources):
    v = np.ascontiguousarray(vertices, dtype=np.uint32)
    e = np.ascontiguousarray(edges, dtype=np.uint64)
    r = np.ascontiguousarray(row_ptr, dtype=np.int32)
    c = np.ascontiguousarray(col, dtype=np.int32)
    s = None if sources is None else np.ascontiguousarray(sources, dtype=np.uint32)
    return v, e, r, c, s


def bfs_levels(vertices, edges, row_ptr, col, sources=None, threads=0):
    """Textbook queue BFS of every graph of a batch."""
    v, e, r, c, s = _bfs_args(vertices, edges, row_ptr, col, sources)
    out = np.zeros(int(v.sum()), dtype=np.int32)
    lib().oracle_bfs_levels(len(v), v.ctypes.data, e.ctypes.data, r.ctypes.data, c.ctypes.data if c.size else None,
                            _ptr(s), out.ctypes.data, threads)
    return out


def bfs_program(vertices, edges, row_ptr, col, sources=None, split=0, threads=1):
    """The BFS specification on the host executors: (levels, stats)."""
    v, e, r, c, s = _bfs_args(vertices, edges, row_ptr, col, sources)
    out = np.zeros(int(v.sum()), dtype=np.int32)
    st = OracleStats()
    lib().oracle_bfs_program(len(v), v.ctypes.data, e.ctypes.data, r.ctypes.data, c.ctypes.data if c.size else None,
                             _ptr(s), out.ctypes.data, split, threads, ctypes.byref(st))
    return out, st
