"""Python handle on libharmonize_amd.so, the gfx950 asynchronous execution runtime.

The runtime itself is C++/HIP behind the C ABI declared in include/harmonize_amd.h;
this module only loads that library with ctypes and wraps its entry points for
tests and benchmarks. There is no CPU fallback: if the library is missing or
cannot be loaded, every call raises.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libharmonize_amd.so"

EVENT = 0
ASYNC = 1
PLAIN = 2
KINDS = {"event": EVENT, "async": ASYNC, "plain": PLAIN}

COLLATZ_OVERFLOW = 0xFFFFFFFF


class HarmonizeError(RuntimeError):
    """A failure reported by the runtime library."""


class RunConfig(ctypes.Structure):
    _fields_ = [
        ("groups", ctypes.c_uint32),
        ("cycles", ctypes.c_uint32),
        ("arena", ctypes.c_uint64),
        ("queues", ctypes.c_uint32),
        ("coalesce", ctypes.c_uint32),
        ("max_launches", ctypes.c_uint32),
        ("block_groups", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("spec_param", ctypes.c_uint32),
        ("stash", ctypes.c_uint32),
    ]


RUN_GRAPH = 1
RUN_AUDIT = 2


class RunStats(ctypes.Structure):
    _fields_ = [
        ("promises", ctypes.c_uint64),
        ("launches", ctypes.c_uint64),
        ("spills", ctypes.c_uint64),
        ("fills", ctypes.c_uint64),
        ("make_work_calls", ctypes.c_uint64),
        ("groups", ctypes.c_uint32),
        ("audits", ctypes.c_uint32),
        ("elapsed_ms", ctypes.c_double),
        ("phase_cycles", ctypes.c_uint64 * 8),
        ("links_run", ctypes.c_uint64),
        ("arena_links", ctypes.c_uint64),
        ("forwarded", ctypes.c_uint64),
    ]

    PHASES = ("cycle", "spill", "pull", "make_work", "pick", "read", "eval", "idle")

    def as_dict(self) -> dict:
        out = {name: getattr(self, name) for name, _ in self._fields_ if name != "phase_cycles"}
        if any(self.phase_cycles):
            out["phase_cycles"] = dict(zip(self.PHASES, [int(x) for x in self.phase_cycles]))
        return out


class DeviceFacts(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("compute_units", ctypes.c_int32),
        ("max_groups_collatz_async", ctypes.c_int32),
        ("max_groups_collatz_event", ctypes.c_int32),
        ("total_memory", ctypes.c_int64),
        ("name", ctypes.c_char * 64),
        ("arch", ctypes.c_char * 64),
    ]


_lib = None

# C signatures of the entry points used from Python: name -> (restype, argtypes).
_VP = ctypes.c_void_p
_SIGNATURES = {
    "hz_last_error": (ctypes.c_char_p, []),
    "harmonize_alloc_device_bytes": (_VP, [ctypes.c_size_t]),
    "harmonize_alloc_managed_bytes": (_VP, [ctypes.c_size_t]),
    "harmonize_free_device_bytes": (None, [_VP]),
    "harmonize_memcpy_host_to_device": (None, [_VP, _VP, ctypes.c_size_t]),
    "harmonize_memcpy_device_to_host": (None, [_VP, _VP, ctypes.c_size_t]),
    "hz_device_synchronize": (ctypes.c_int, []),
    "hz_device_copy": (ctypes.c_int, [_VP, _VP, ctypes.c_size_t]),
    "hz_device_query": (ctypes.c_int, [ctypes.POINTER(DeviceFacts)]),
    "hz_collatz_run": (
        ctypes.c_int,
        [ctypes.c_int, _VP, ctypes.c_uint32, _VP, ctypes.POINTER(RunConfig), ctypes.POINTER(RunStats)],
    ),
    "hz_collatz_make_starts": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _VP]),
    "hz_audit_selftest": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "hz_tune_cache_set_path": (ctypes.c_int, [ctypes.c_char_p]),
    "hz_tune_cache_lookup": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(RunConfig),
                                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]),
    "hz_tune_cache_store": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(RunConfig), ctypes.c_double]),
    "hz_chase_run": (
        ctypes.c_int,
        [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, _VP, _VP, _VP, _VP, ctypes.POINTER(RunConfig),
         ctypes.POINTER(RunStats)],
    ),
    "hz_tree_run": (
        ctypes.c_int,
        [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _VP, _VP,
         ctypes.POINTER(RunConfig), ctypes.POINTER(RunStats)],
    ),
    "hz_collatz_run_noalloc": (
        ctypes.c_int, [_VP, ctypes.c_uint32, _VP, ctypes.POINTER(RunConfig), ctypes.POINTER(RunStats)],
    ),
}

PROGRAM_SUFFIXES = ("collatz_event", "collatz_async", "photon_event", "photon_async", "bfs_event", "bfs_async",
                    "barrier_async")
for _s in PROGRAM_SUFFIXES:
    _SIGNATURES.update(
        {
            f"alloc_state_{_s}": (_VP, []),
            f"free_state_{_s}": (None, [_VP]),
            f"load_state_device_{_s}": (None, [_VP, _VP]),
            f"store_state_device_{_s}": (None, [_VP, _VP]),
            f"alloc_program_{_s}": (_VP, [_VP, ctypes.c_size_t]),
            f"free_program_{_s}": (None, [_VP]),
            f"init_program_{_s}": (None, [_VP, ctypes.c_size_t]),
            f"exec_program_{_s}": (None, [_VP, ctypes.c_size_t, ctypes.c_size_t]),
            f"complete_{_s}": (ctypes.c_int, [_VP]),
            f"clear_flags_{_s}": (None, [_VP]),
            f"set_device_{_s}": (None, [ctypes.c_int]),
        }
    )


def lib() -> ctypes.CDLL:
    """The loaded runtime library (raises HarmonizeError if it is not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise HarmonizeError(
                f"{LIB_PATH} is not built; run `python -m harmonize_amd.build` (hipcc, gfx950)"
            )
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def last_error() -> str | None:
    msg = lib().hz_last_error()
    return msg.decode() if msg else None


def check(status: int, what: str = "call") -> None:
    if status != 0:
        raise HarmonizeError(f"{what} failed ({status}): {last_error()}")


def check_void(what: str) -> None:
    """For the reference-shaped void entry points: raise if they recorded an error."""
    msg = last_error()
    if msg:
        raise HarmonizeError(f"{what} failed: {msg}")


class DeviceArray:
    """A device allocation holding a 1-D numpy-typed array."""

    def __init__(self, count: int, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = max(1, self.count * self.dtype.itemsize)
        ptr = lib().harmonize_alloc_device_bytes(self.nbytes)
        if not ptr:
            raise HarmonizeError(f"device allocation of {self.nbytes} bytes failed: {last_error()}")
        self.ptr = ptr

    @classmethod
    def from_numpy(cls, host: np.ndarray) -> "DeviceArray":
        host = np.ascontiguousarray(host)
        arr = cls(host.size, host.dtype)
        arr.upload(host)
        return arr

    def upload(self, host: np.ndarray) -> None:
        host = np.ascontiguousarray(host, dtype=self.dtype)
        if host.size != self.count:
            raise ValueError("size mismatch")
        if self.count:
            lib().harmonize_memcpy_host_to_device(self.ptr, host.ctypes.data, host.nbytes)
            check_void("upload")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.count, dtype=self.dtype)
        if self.count:
            lib().harmonize_memcpy_device_to_host(out.ctypes.data, self.ptr, out.nbytes)
            check_void("download")
        return out

    def free(self) -> None:
        if getattr(self, "ptr", None):
            lib().harmonize_free_device_bytes(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_copy(dst: "DeviceArray", src: "DeviceArray", nbytes: int) -> None:
    """Synchronous device-to-device copy of nbytes between two device arrays."""
    if nbytes > min(dst.nbytes, src.nbytes):
        raise ValueError("device_copy: nbytes exceeds an array")
    check(lib().hz_device_copy(dst.ptr, src.ptr, nbytes), "hz_device_copy")


def device_facts() -> dict:
    f = DeviceFacts()
    check(lib().hz_device_query(ctypes.byref(f)), "hz_device_query")
    return {
        "device": f.device,
        "compute_units": f.compute_units,
        "max_groups_collatz_async": f.max_groups_collatz_async,
        "max_groups_collatz_event": f.max_groups_collatz_event,
        "total_memory": f.total_memory,
        "name": f.name.decode(),
        "arch": f.arch.decode(),
    }


def make_config(**kw) -> RunConfig:
    cfg = RunConfig()
    for k, v in kw.items():
        setattr(cfg, k, int(v))
    return cfg


def collatz_run(kind, steps: DeviceArray, starts: DeviceArray | None = None, count: int | None = None, **cfg) -> dict:
    """Runs the Collatz program on the GPU; returns its statistics."""
    k = KINDS[kind] if isinstance(kind, str) else int(kind)
    n = steps.count if count is None else int(count)
    st = RunStats()
    config = make_config(**cfg)
    check(
        lib().hz_collatz_run(k, starts.ptr if starts is not None else None, n, steps.ptr, ctypes.byref(config),
                             ctypes.byref(st)),
        "hz_collatz_run",
    )
    return st.as_dict()


def collatz_run_noalloc(steps: DeviceArray, starts: DeviceArray | None = None, count: int | None = None,
                        **cfg) -> dict:
    """Runs the Collatz AsyncProgram with the runtime's buffers in the
    library's fixed device array (hz_collatz_run_noalloc); returns its statistics."""
    n = steps.count if count is None else int(count)
    st = RunStats()
    config = make_config(**cfg)
    check(
        lib().hz_collatz_run_noalloc(starts.ptr if starts is not None else None, n, steps.ptr, ctypes.byref(config),
                                     ctypes.byref(st)),
        "hz_collatz_run_noalloc",
    )
    return st.as_dict()


def collatz_make_starts(seed: int, bits: int, count: int) -> DeviceArray:
    out = DeviceArray(count, np.uint64)
    check(lib().hz_collatz_make_starts(seed, bits, count, out.ptr), "hz_collatz_make_starts")
    return out


# ------------------------------------------------------------------ photon walk

PHOTON_EXTRA_SLOTS = 5
PHOTON_SLOTS = ("reflected", "transmitted", "roulette_lost", "roulette_gained", "capped")


class PhotonParams(ctypes.Structure):
    _fields_ = [
        ("photons", ctypes.c_uint32),
        ("bins", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("thickness", ctypes.c_float),
        ("sigma_t", ctypes.c_float),
        ("albedo", ctypes.c_float),
        ("split_above", ctypes.c_float),
        ("roulette_below", ctypes.c_float),
        ("region_width", ctypes.c_float),
        ("regions", ctypes.c_uint32),
        ("max_split", ctypes.c_uint32),
        ("max_events", ctypes.c_uint32),
        ("replicas", ctypes.c_uint32),
    ]


_SIGNATURES.update(
    {
        "hz_photon_run": (
            ctypes.c_int,
            [ctypes.c_int, ctypes.POINTER(PhotonParams), _VP, _VP, ctypes.POINTER(RunConfig), ctypes.POINTER(RunStats)],
        ),
        "hz_photon_defaults": (None, [ctypes.POINTER(PhotonParams)]),
        "hz_detmath_eval": (ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_uint32]),
        "hz_barrier_run": (
            ctypes.c_int,
            [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _VP, _VP, _VP, ctypes.POINTER(RunConfig),
             ctypes.POINTER(RunStats)],
        ),
        "hz_futures_run": (
            ctypes.c_int,
            [ctypes.c_int, _VP, ctypes.c_uint32, ctypes.c_uint32, _VP, _VP, _VP, ctypes.POINTER(RunConfig),
             ctypes.POINTER(RunStats)],
        ),
    }
)

def barrier_run(dep_count: int, non_dep_count: int, barrier_count: int, **cfg):
    """Runs the remapping-barrier program; returns (counter, hits[dep], violations, stats)."""
    counter = DeviceArray(1, np.uint32)
    violations = DeviceArray(1, np.uint32)
    hits = DeviceArray(max(1, dep_count), np.uint32)
    st = RunStats()
    config = make_config(**cfg)
    check(lib().hz_barrier_run(dep_count, non_dep_count, barrier_count, counter.ptr, hits.ptr, violations.ptr,
                               ctypes.byref(config), ctypes.byref(st)), "hz_barrier_run")
    return int(counter.numpy()[0]), hits.numpy()[:dep_count], int(violations.numpy()[0]), st.as_dict()


TREE_FLAT, TREE_DEEPEST, TREE_SHALLOWEST = 0, 1, 2


def chase_run(kind, lazy: bool, heads: DeviceArray, nodes: DeviceArray, **cfg):
    """Runs the pointer-chasing program over lists starting at `heads` (u32)
    through `nodes` (u32 pairs [next, value]); returns (sums, lengths, stats)."""
    k = KINDS[kind] if isinstance(kind, str) else int(kind)
    lists = heads.count
    sums = DeviceArray(max(1, lists), np.uint64)
    lengths = DeviceArray(max(1, lists), np.uint32)
    st = RunStats()
    config = make_config(**cfg)
    check(lib().hz_chase_run(k, int(bool(lazy)), lists, heads.ptr, nodes.ptr, sums.ptr, lengths.ptr,
                             ctypes.byref(config), ctypes.byref(st)), "hz_chase_run")
    return sums.numpy()[:lists], lengths.numpy()[:lists], st.as_dict()


def tree_run(kind, order: int, roots: int, fanout: int, height: int, **cfg):
    """Runs the depth-level tree program; returns (leaves[roots], violations, stats)."""
    k = KINDS[kind] if isinstance(kind, str) else int(kind)
    leaves = DeviceArray(max(1, roots), np.uint32)
    violations = DeviceArray(1, np.uint32)
    st = RunStats()
    config = make_config(**cfg)
    check(lib().hz_tree_run(k, order, roots, fanout, height, leaves.ptr, violations.ptr, ctypes.byref(config),
                            ctypes.byref(st)), "hz_tree_run")
    return leaves.numpy()[:roots], int(violations.numpy()[0]), st.as_dict()


FUTURES_UNRESOLVED = (1 << 64) - 1  # a reader that ran early, or a future left unresolved


def futures_run(kind, starts: DeviceArray, readers: int = 0, **cfg):
    """Runs the futures program (value-returning operations resolving a tree of
    futures over the Collatz step counts of `starts`, a power-of-two count).
    Returns (total, echo[readers], forwarded[readers], stats): total is the sum
    of the step counts; reader r writes echo[r] = total + r after awaiting the
    total and receives forwarded[r] = total through a forwarded future."""
    k = KINDS[kind] if isinstance(kind, str) else int(kind)
    echo = DeviceArray(max(1, readers), np.uint64)
    forwarded = DeviceArray(max(1, readers), np.uint64)
    total = ctypes.c_uint64(0)
    st = RunStats()
    config = make_config(**cfg)
    check(lib().hz_futures_run(k, starts.ptr, starts.count, readers, echo.ptr, forwarded.ptr, ctypes.byref(total),
                               ctypes.byref(config), ctypes.byref(st)), "hz_futures_run")
    return int(total.value), echo.numpy()[:readers], forwarded.numpy()[:readers], st.as_dict()


def photon_params(**overrides) -> PhotonParams:
    """The photon-walk defaults (1e7 photons, slab 20, albedo 0.9, ...) with overrides."""
    p = PhotonParams()
    lib().hz_photon_defaults(ctypes.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def photon_run(kind, params: PhotonParams, tally: DeviceArray | None = None, events: DeviceArray | None = None,
               **cfg):
    """Runs the photon walk; returns (summed tally, events per photon, stats)."""
    k = KINDS[kind] if isinstance(kind, str) else int(kind)
    stride = params.bins + PHOTON_EXTRA_SLOTS
    if tally is None:
        tally = DeviceArray(params.replicas * stride, np.float64)
    if events is None:
        events = DeviceArray(params.photons, np.uint32)
    st = RunStats()
    config = make_config(**cfg)
    check(lib().hz_photon_run(k, ctypes.byref(params), tally.ptr, events.ptr, ctypes.byref(config), ctypes.byref(st)),
          "hz_photon_run")
    reps = tally.numpy().reshape(params.replicas, stride)
    return reps.sum(axis=0), events.numpy(), st.as_dict()


# ------------------------------------------------------------------ batched BFS

BFS_UNREACHED = 0x7FFFFFFF


class BfsInfo(ctypes.Structure):
    _fields_ = [
        ("graphs", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("total_vertices", ctypes.c_uint64),
        ("total_edges", ctypes.c_uint64),
        ("d_row_ptr", _VP),
        ("d_col", _VP),
        ("d_level", _VP),
        ("build_ms", ctypes.c_double),
    ]


_SIGNATURES.update(
    {
        "hz_bfs_batch_sizes": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double, ctypes.c_double, _VP]),
        "hz_bfs_batch_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, _VP, ctypes.c_uint32,
                                               ctypes.POINTER(_VP)]),
        "hz_bfs_batch_from_csr": (ctypes.c_int, [ctypes.c_uint32, _VP, _VP, _VP, _VP, _VP, ctypes.POINTER(_VP)]),
        "hz_bfs_batch_destroy": (ctypes.c_int, [_VP]),
        "hz_bfs_batch_info": (ctypes.c_int, [_VP, ctypes.POINTER(BfsInfo)]),
        "hz_bfs_batch_layout": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, _VP]),
        "hz_bfs_batch_run": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.POINTER(RunConfig), ctypes.POINTER(RunStats)]),
        "hz_bfs_batch_release": (ctypes.c_int, [_VP]),
    }
)


def bfs_batch_sizes(seed: int, graphs: int, min_log2: float = 14.0, max_log2: float = 22.0) -> np.ndarray:
    """Vertex counts of an R-MAT batch (host computation in the library)."""
    out = np.zeros(graphs, dtype=np.uint32)
    check(lib().hz_bfs_batch_sizes(seed, graphs, min_log2, max_log2, out.ctypes.data), "hz_bfs_batch_sizes")
    return out


class BfsBatch:
    """Many CSR graphs resident on the device, searched by one program per run."""

    def __init__(self, handle):
        self.handle = handle
        info = BfsInfo()
        check(lib().hz_bfs_batch_info(handle, ctypes.byref(info)), "hz_bfs_batch_info")
        self.graphs = info.graphs
        self.total_vertices = info.total_vertices
        self.total_edges = info.total_edges
        self.build_ms = info.build_ms
        self._d_row_ptr, self._d_col, self._d_level = info.d_row_ptr, info.d_col, info.d_level
        g = self.graphs
        self.vertices = np.zeros(g, np.uint32)
        self.edges = np.zeros(g, np.uint64)
        self.vertex_base = np.zeros(g, np.uint64)
        self.edge_base = np.zeros(g, np.uint64)
        self.row_base = np.zeros(g, np.uint64)
        check(lib().hz_bfs_batch_layout(handle, self.vertices.ctypes.data, self.edges.ctypes.data,
                                        self.vertex_base.ctypes.data, self.edge_base.ctypes.data,
                                        self.row_base.ctypes.data), "hz_bfs_batch_layout")

    @classmethod
    def rmat(cls, seed: int, vertices, edge_factor: int = 16) -> "BfsBatch":
        v = np.ascontiguousarray(vertices, dtype=np.uint32)
        h = _VP()
        check(lib().hz_bfs_batch_create(seed, len(v), v.ctypes.data, edge_factor, ctypes.byref(h)),
              "hz_bfs_batch_create")
        return cls(h.value)

    @classmethod
    def from_csr(cls, vertices, edges, row_ptr, col, sources=None) -> "BfsBatch":
        v = np.ascontiguousarray(vertices, dtype=np.uint32)
        e = np.ascontiguousarray(edges, dtype=np.uint64)
        r = np.ascontiguousarray(row_ptr, dtype=np.int32)
        c = np.ascontiguousarray(col, dtype=np.int32)
        s = None if sources is None else np.ascontiguousarray(sources, dtype=np.uint32)
        h = _VP()
        check(lib().hz_bfs_batch_from_csr(len(v), v.ctypes.data, e.ctypes.data, r.ctypes.data,
                                          c.ctypes.data if c.size else None,
                                          None if s is None else s.ctypes.data, ctypes.byref(h)),
              "hz_bfs_batch_from_csr")
        return cls(h.value)

    def run(self, kind="async", **cfg) -> dict:
        k = KINDS[kind] if isinstance(kind, str) else int(kind)
        st = RunStats()
        config = make_config(**cfg)
        check(lib().hz_bfs_batch_run(self.handle, k, ctypes.byref(config), ctypes.byref(st)), "hz_bfs_batch_run")
        return st.as_dict()

    def _download(self, ptr, count, dtype):
        out = np.empty(int(count), dtype=dtype)
        if out.size:
            lib().harmonize_memcpy_device_to_host(out.ctypes.data, ptr, out.nbytes)
            check_void("download")
        return out

    def levels(self) -> np.ndarray:
        return self._download(self._d_level, self.total_vertices, np.int32)

    def csr(self):
        """(row_ptr, col) of the whole batch, as laid out on the device."""
        return (self._download(self._d_row_ptr, self.total_vertices + self.graphs, np.int32),
                self._download(self._d_col, self.total_edges, np.int32))

    def csr_rows(self):
        """(row_ptr, None): the row offsets alone, without the column array."""
        return self._download(self._d_row_ptr, self.total_vertices + self.graphs, np.int32), None

    def release(self) -> None:
        check(lib().hz_bfs_batch_release(self.handle), "hz_bfs_batch_release")

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().hz_bfs_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ autotuning


class TuneResult(ctypes.Structure):
    _fields_ = [
        ("best", RunConfig),
        ("best_ms", ctypes.c_double),
        ("start_ms", ctypes.c_double),
        ("trials", ctypes.c_uint32),
        ("failures", ctypes.c_uint32),
        ("cached", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class TuneTrial(ctypes.Structure):
    _fields_ = [("config", RunConfig), ("ms", ctypes.c_double)]


_TUNE_TAIL = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(TuneResult), ctypes.POINTER(TuneTrial), ctypes.c_uint32]
_SIGNATURES.update(
    {
        "hz_collatz_tune": (ctypes.c_int, [_VP, ctypes.c_uint32, _VP] + _TUNE_TAIL),
        "hz_photon_tune": (ctypes.c_int, [ctypes.POINTER(PhotonParams), _VP, _VP] + _TUNE_TAIL),
        "hz_bfs_batch_tune": (ctypes.c_int, [_VP] + _TUNE_TAIL),
    }
)


def _config_dict(c: RunConfig) -> dict:
    return {name: int(getattr(c, name)) for name, _ in RunConfig._fields_}


def _tune_out(res: TuneResult, log, n: int) -> dict:
    return {
        "best": _config_dict(res.best),
        "best_ms": res.best_ms,
        "start_ms": res.start_ms,
        "trials": res.trials,
        "failures": res.failures,
        "cached": bool(res.cached),
        "log": [(_config_dict(log[i].config), log[i].ms) for i in range(min(n, res.trials))],
    }


def tune_cache_set_path(path: str | None) -> None:
    """Sets the tuning cache file (None disables it); see hz_tune_cache_set_path."""
    check(lib().hz_tune_cache_set_path(path.encode() if path else None), "hz_tune_cache_set_path")


def tune_cache_lookup(key: str):
    """(config dict, best_ms) stored under `key`, or None."""
    c, ms, found = RunConfig(), ctypes.c_double(), ctypes.c_int()
    check(lib().hz_tune_cache_lookup(key.encode(), ctypes.byref(c), ctypes.byref(ms), ctypes.byref(found)),
          "hz_tune_cache_lookup")
    return (_config_dict(c), ms.value) if found.value else None


def tune_cache_store(key: str, config: dict, best_ms: float) -> None:
    check(lib().hz_tune_cache_store(key.encode(), ctypes.byref(make_config(**config)), best_ms),
          "hz_tune_cache_store")


def tune_collatz(steps: DeviceArray, starts: DeviceArray | None, count: int, rounds: int = 2, reps: int = 1) -> dict:
    res, log = TuneResult(), (TuneTrial * 256)()
    check(lib().hz_collatz_tune(starts.ptr if starts is not None else None, count, steps.ptr, rounds, reps,
                                ctypes.byref(res), log, 256), "hz_collatz_tune")
    return _tune_out(res, log, 256)


def tune_photon(params: PhotonParams, rounds: int = 2, reps: int = 1) -> dict:
    tally = DeviceArray(params.replicas * (params.bins + PHOTON_EXTRA_SLOTS), np.float64)
    events = DeviceArray(params.photons, np.uint32)
    res, log = TuneResult(), (TuneTrial * 256)()
    check(lib().hz_photon_tune(ctypes.byref(params), tally.ptr, events.ptr, rounds, reps, ctypes.byref(res), log, 256),
          "hz_photon_tune")
    return _tune_out(res, log, 256)


def tune_bfs(batch: "BfsBatch", rounds: int = 2, reps: int = 1) -> dict:
    res, log = TuneResult(), (TuneTrial * 256)()
    check(lib().hz_bfs_batch_tune(batch.handle, rounds, reps, ctypes.byref(res), log, 256), "hz_bfs_batch_tune")
    return _tune_out(res, log, 256)


# ------------------------------------------------------------------ serving

JOB_COLLATZ, JOB_PHOTON, JOB_BFS = 0, 1, 2
JOB_TYPES = {"collatz": JOB_COLLATZ, "photon": JOB_PHOTON, "bfs": JOB_BFS}
JOB_PENDING, JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2, -1


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("pool_bytes", ctypes.c_uint64),
        ("arena_links", ctypes.c_uint64),
        ("slots", ctypes.c_uint32),
        ("max_pending", ctypes.c_uint32),
        ("keep_outputs", ctypes.c_uint32),
        ("policy", ctypes.c_uint32),
    ]


SERVE_FIFO, SERVE_LARGEST_FIRST = 0, 1


class Job(ctypes.Structure):
    _fields_ = [("type", ctypes.c_uint32), ("size", ctypes.c_uint32), ("seed", ctypes.c_uint64)]


class JobResult(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64),
        ("type", ctypes.c_uint32),
        ("size", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("status", ctypes.c_int32),
        ("slot", ctypes.c_uint32),
        ("promises", ctypes.c_uint64),
        ("launches", ctypes.c_uint32),
        ("groups", ctypes.c_uint32),
        ("submit_ms", ctypes.c_double),
        ("start_ms", ctypes.c_double),
        ("end_ms", ctypes.c_double),
        ("state_bytes", ctypes.c_uint64),
        ("chunks", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class ServerInfo(ctypes.Structure):
    _fields_ = [
        ("slots", ctypes.c_uint32),
        ("pending", ctypes.c_uint32),
        ("jobs", ctypes.c_uint64),
        ("pool_bytes", ctypes.c_uint64),
        ("pool_used", ctypes.c_uint64),
        ("pool_largest_free", ctypes.c_uint64),
        ("groups_collatz", ctypes.c_uint32),
        ("groups_photon", ctypes.c_uint32),
        ("groups_bfs", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


_SIGNATURES.update(
    {
        "hz_server_create": (ctypes.c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_VP)]),
        "hz_server_destroy": (ctypes.c_int, [_VP]),
        "hz_server_submit": (ctypes.c_int, [_VP, ctypes.POINTER(Job), ctypes.POINTER(ctypes.c_uint64)]),
        "hz_server_progress": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]),
        "hz_server_drain": (ctypes.c_int, [_VP, ctypes.c_double]),
        "hz_server_job": (ctypes.c_int, [_VP, ctypes.c_uint64, ctypes.POINTER(JobResult)]),
        "hz_server_output": (ctypes.c_int, [_VP, ctypes.c_uint64, ctypes.c_uint32, _VP, ctypes.c_uint64]),
        "hz_server_release": (ctypes.c_int, [_VP, ctypes.c_uint64]),
        "hz_server_info": (ctypes.c_int, [_VP, ctypes.POINTER(ServerInfo)]),
        "hz_server_run_trace": (ctypes.c_int, [_VP, ctypes.POINTER(Job), _VP, ctypes.c_uint32, ctypes.c_double,
                                               ctypes.POINTER(JobResult)]),
        "hz_pool_test": (ctypes.c_int, [ctypes.c_uint64, _VP, ctypes.c_uint32, _VP]),
    }
)


def pool_test(size: int, ops) -> np.ndarray:
    """Runs the server's job-state allocator on the host (no GPU): ops >= 0
    allocate, -k-1 frees allocation k. Returns offsets / largest free blocks."""
    o = np.ascontiguousarray(ops, dtype=np.int64)
    out = np.zeros(len(o), dtype=np.uint64)
    check(lib().hz_pool_test(size, o.ctypes.data, len(o), out.ctypes.data), "hz_pool_test")
    return out


def _job(kind, size: int, seed: int) -> Job:
    t = JOB_TYPES[kind] if isinstance(kind, str) else int(kind)
    return Job(t, int(size), int(seed))


class Server:
    """A request server running Collatz, photon and BFS jobs side by side."""

    OUTPUTS = {
        JOB_COLLATZ: [(np.uint32, 1), (np.uint64, 1)],
        JOB_PHOTON: [(np.uint32, 1), (np.float64, None)],
        JOB_BFS: [(np.int32, 1), (np.int32, 1), (np.int32, 16)],
    }

    def __init__(self, pool_bytes: int = 0, slots: int = 0, arena_links: int = 0, keep_outputs: bool = False,
                 max_pending: int = 0, policy: int = SERVE_FIFO):
        cfg = ServerConfig(pool_bytes, arena_links, slots, max_pending, 1 if keep_outputs else 0, policy)
        h = _VP()
        check(lib().hz_server_create(ctypes.byref(cfg), ctypes.byref(h)), "hz_server_create")
        self.handle = h.value

    def submit(self, kind, size: int, seed: int) -> int:
        i = ctypes.c_uint64()
        j = _job(kind, size, seed)
        check(lib().hz_server_submit(self.handle, ctypes.byref(j), ctypes.byref(i)), "hz_server_submit")
        return i.value

    def progress(self):
        r, p = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().hz_server_progress(self.handle, ctypes.byref(r), ctypes.byref(p)), "hz_server_progress")
        return r.value, p.value

    def drain(self, timeout_ms: float = 600000.0) -> None:
        check(lib().hz_server_drain(self.handle, timeout_ms), "hz_server_drain")

    def job(self, i: int) -> dict:
        r = JobResult()
        check(lib().hz_server_job(self.handle, i, ctypes.byref(r)), "hz_server_job")
        return r.as_dict()

    def output(self, i: int, which: int = 0) -> np.ndarray:
        info = self.job(i)
        dtype, per = self.OUTPUTS[info["type"]][which]
        if info["type"] == JOB_PHOTON and which == 1:
            count = info["chunks"] * 64 * (1000 + PHOTON_EXTRA_SLOTS)
        elif info["type"] == JOB_BFS and which == 1:
            count = info["size"] + 1
        else:
            count = info["size"] * per
        out = np.empty(count, dtype=dtype)
        check(lib().hz_server_output(self.handle, i, which, out.ctypes.data, out.nbytes), "hz_server_output")
        return out

    def release(self, i: int) -> None:
        check(lib().hz_server_release(self.handle, i), "hz_server_release")

    def info(self) -> dict:
        s = ServerInfo()
        check(lib().hz_server_info(self.handle, ctypes.byref(s)), "hz_server_info")
        return {name: getattr(s, name) for name, _ in s._fields_ if name != "reserved"}

    def run_trace(self, jobs, arrival_ms, timeout_ms: float = 600000.0) -> list:
        """Replays (kind, size, seed) jobs arriving at arrival_ms (sorted, ms
        after the call); returns one result dict per job."""
        n = len(jobs)
        arr = (Job * n)(*[_job(*j) for j in jobs])
        t = np.ascontiguousarray(arrival_ms, dtype=np.float64)
        res = (JobResult * n)()
        check(lib().hz_server_run_trace(self.handle, arr, t.ctypes.data, n, timeout_ms, res), "hz_server_run_trace")
        return [r.as_dict() for r in res]

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().hz_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ resident serving

RESIDENT_BY_ADMISSION = 0
RESIDENT_SMALLEST_FIRST = 1


class ResidentConfig(ctypes.Structure):
    _fields_ = [
        ("pool_bytes", ctypes.c_uint64),
        ("arena_links", ctypes.c_uint64),
        ("groups", ctypes.c_uint32),
        ("policy", ctypes.c_uint32),
        ("idle_us", ctypes.c_uint32),
        ("poll_us", ctypes.c_uint32),
        ("life_ms", ctypes.c_uint32),
        ("bits", ctypes.c_uint32),
        ("keep_outputs", ctypes.c_uint32),
        ("kind", ctypes.c_uint32),
    ]


class ResidentInfo(ctypes.Structure):
    _fields_ = [
        ("jobs", ctypes.c_uint64),
        ("pending", ctypes.c_uint32),
        ("in_flight", ctypes.c_uint32),
        ("launches", ctypes.c_uint32),
        ("joined_running", ctypes.c_uint32),
        ("groups", ctypes.c_uint32),
        ("running", ctypes.c_uint32),
        ("pool_bytes", ctypes.c_uint64),
        ("pool_used", ctypes.c_uint64),
        ("promises", ctypes.c_uint64),
        ("links_run", ctypes.c_uint64),
        ("make_work_calls", ctypes.c_uint64),
        ("idle_polls", ctypes.c_uint64),
        ("phase_cycles", ctypes.c_uint64 * 8),
    ]


_SIGNATURES.update(
    {
        "hz_resident_create": (ctypes.c_int, [ctypes.POINTER(ResidentConfig), ctypes.POINTER(_VP)]),
        "hz_resident_destroy": (ctypes.c_int, [_VP]),
        "hz_resident_submit": (ctypes.c_int, [_VP, ctypes.POINTER(Job), ctypes.POINTER(ctypes.c_uint64)]),
        "hz_resident_poll": (ctypes.c_int, [_VP, _VP, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
        "hz_resident_drain": (ctypes.c_int, [_VP, ctypes.c_double]),
        "hz_resident_job": (ctypes.c_int, [_VP, ctypes.c_uint64, ctypes.POINTER(JobResult)]),
        "hz_resident_output": (ctypes.c_int, [_VP, ctypes.c_uint64, ctypes.c_uint32, _VP, ctypes.c_uint64]),
        "hz_resident_release": (ctypes.c_int, [_VP, ctypes.c_uint64]),
        "hz_resident_info": (ctypes.c_int, [_VP, ctypes.POINTER(ResidentInfo)]),
        "hz_resident_run_trace": (ctypes.c_int, [_VP, ctypes.POINTER(Job), _VP, ctypes.c_uint32, ctypes.c_double,
                                                 ctypes.POINTER(JobResult)]),
    }
)


class Resident:
    """Collatz or photon-walk jobs admitted into a running persistent kernel
    (hz_resident_*); `kind` is "collatz" or "photon"."""

    def __init__(self, pool_bytes: int = 0, arena_links: int = 0, groups: int = 0,
                 policy: int = RESIDENT_BY_ADMISSION, idle_us: int = 0, poll_us: int = 0, life_ms: int = 0,
                 bits: int = 0, keep_outputs: bool = False, kind: str = "collatz"):
        self.kind = kind
        cfg = ResidentConfig(pool_bytes, arena_links, groups, policy, idle_us, poll_us, life_ms, bits,
                             1 if keep_outputs else 0, JOB_TYPES[kind])
        h = _VP()
        check(lib().hz_resident_create(ctypes.byref(cfg), ctypes.byref(h)), "hz_resident_create")
        self.handle = h.value

    def submit(self, size: int, seed: int) -> int:
        i = ctypes.c_uint64()
        j = _job(self.kind, size, seed)
        check(lib().hz_resident_submit(self.handle, ctypes.byref(j), ctypes.byref(i)), "hz_resident_submit")
        return i.value

    def poll(self, capacity: int = 64) -> list:
        ids = np.zeros(capacity, dtype=np.uint64)
        n = ctypes.c_uint32()
        check(lib().hz_resident_poll(self.handle, ids.ctypes.data, capacity, ctypes.byref(n)), "hz_resident_poll")
        return [int(x) for x in ids[: n.value]]

    def drain(self, timeout_ms: float = 600000.0) -> None:
        check(lib().hz_resident_drain(self.handle, timeout_ms), "hz_resident_drain")

    def job(self, i: int) -> dict:
        r = JobResult()
        check(lib().hz_resident_job(self.handle, i, ctypes.byref(r)), "hz_resident_job")
        return r.as_dict()

    def output(self, i: int, which: int = 0) -> np.ndarray:
        n = self.job(i)["size"]
        if which == 0:
            out = np.empty(n, dtype=np.uint32)
        elif self.kind == "photon":
            out = np.empty(64 * (1000 + PHOTON_EXTRA_SLOTS), dtype=np.float64)
        else:
            out = np.empty(n, dtype=np.uint64)
        check(lib().hz_resident_output(self.handle, i, which, out.ctypes.data, out.nbytes), "hz_resident_output")
        return out

    def release(self, i: int) -> None:
        check(lib().hz_resident_release(self.handle, i), "hz_resident_release")

    def info(self) -> dict:
        s = ResidentInfo()
        check(lib().hz_resident_info(self.handle, ctypes.byref(s)), "hz_resident_info")
        out = {name: getattr(s, name) for name, _ in s._fields_}
        out["phase_cycles"] = list(s.phase_cycles)
        return out

    def run_trace(self, sizes_seeds, arrival_ms, timeout_ms: float = 600000.0) -> list:
        """Replays (size, seed) jobs of the server's kind arriving at
        arrival_ms (sorted, ms after the call); returns one result dict per
        job."""
        n = len(sizes_seeds)
        arr = (Job * n)(*[_job(self.kind, s, sd) for s, sd in sizes_seeds])
        t = np.ascontiguousarray(arrival_ms, dtype=np.float64)
        res = (JobResult * n)()
        check(lib().hz_resident_run_trace(self.handle, arr, t.ctypes.data, n, timeout_ms, res),
              "hz_resident_run_trace")
        return [r.as_dict() for r in res]

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().hz_resident_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ generic instances

SPEC_COLLATZ, SPEC_PHOTON, SPEC_BFS, SPEC_BARRIER = 0, 1, 2, 3

_SIGNATURES.update(
    {
        "hz_instance_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _VP, ctypes.c_size_t,
                                              ctypes.POINTER(_VP)]),
        "hz_instance_destroy": (ctypes.c_int, [_VP]),
        "hz_init": (ctypes.c_int, [_VP, ctypes.c_uint32]),
        "hz_exec": (ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.c_uint32]),
        "hz_complete": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int)]),
        "hz_clear_flags": (ctypes.c_int, [_VP]),
        "hz_state_load": (ctypes.c_int, [_VP, ctypes.c_uint64, ctypes.c_uint64, _VP]),
        "hz_state_store": (ctypes.c_int, [_VP, ctypes.c_uint64, ctypes.c_uint64, _VP]),
        "hz_set_device": (ctypes.c_int, [ctypes.c_int]),
        "hz_init_on": (ctypes.c_int, [_VP, ctypes.c_uint32, _VP]),
        "hz_exec_on": (ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.c_uint32, _VP]),
        "hz_complete_on": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(ctypes.c_int)]),
        "hz_stream_create": (ctypes.c_int, [ctypes.POINTER(_VP)]),
        "hz_stream_destroy": (ctypes.c_int, [_VP]),
        "hz_serve_submit": (ctypes.c_int, [_VP, ctypes.POINTER(Job), ctypes.POINTER(ctypes.c_uint64)]),
        "hz_serve_poll": (ctypes.c_int, [_VP, _VP, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
    }
)


class CollatzState(ctypes.Structure):
    _fields_ = [("starts", _VP), ("steps", _VP), ("iterator", _VP)]


def stream_create():
    """A new non-blocking HIP stream (an opaque handle)."""
    s = _VP()
    check(lib().hz_stream_create(ctypes.byref(s)), "hz_stream_create")
    return s.value


def stream_destroy(stream) -> None:
    check(lib().hz_stream_destroy(stream), "hz_stream_destroy")


class Instance:
    """A program instance chosen at run time: (kind, spec) with a state struct."""

    def __init__(self, kind, spec: int, size: int, state: ctypes.Structure):
        k = KINDS[kind] if isinstance(kind, str) else int(kind)
        h = _VP()
        check(lib().hz_instance_create(k, spec, size, ctypes.byref(state), ctypes.sizeof(state), ctypes.byref(h)),
              "hz_instance_create")
        self.handle = h.value
        self.state_type = type(state)

    def init(self, groups: int) -> None:
        check(lib().hz_init(self.handle, groups), "hz_init")

    def exec(self, groups: int, cycles: int) -> None:
        check(lib().hz_exec(self.handle, groups, cycles), "hz_exec")

    def complete(self) -> bool:
        d = ctypes.c_int()
        check(lib().hz_complete(self.handle, ctypes.byref(d)), "hz_complete")
        return d.value == 1

    def init_on(self, groups: int, stream) -> None:
        """Enqueues init on `stream` (from stream_create) without waiting."""
        check(lib().hz_init_on(self.handle, groups, stream), "hz_init_on")

    def exec_on(self, groups: int, cycles: int, stream) -> None:
        """Enqueues one launch on `stream` without waiting."""
        check(lib().hz_exec_on(self.handle, groups, cycles, stream), "hz_exec_on")

    def complete_on(self, stream) -> bool:
        """Waits for `stream`, then reports completion."""
        d = ctypes.c_int()
        check(lib().hz_complete_on(self.handle, stream, ctypes.byref(d)), "hz_complete_on")
        return d.value == 1

    def load_state(self):
        st = self.state_type()
        check(lib().hz_state_load(self.handle, 0, ctypes.sizeof(st), ctypes.byref(st)), "hz_state_load")
        return st

    def store_state(self, st) -> None:
        check(lib().hz_state_store(self.handle, 0, ctypes.sizeof(st), ctypes.byref(st)), "hz_state_store")

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().hz_instance_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve_poll(server: "Server", capacity: int = 64) -> list:
    """Ids of the server's jobs that finished since the previous poll."""
    ids = np.zeros(capacity, dtype=np.uint64)
    n = ctypes.c_uint32()
    check(lib().hz_serve_poll(server.handle, ids.ctypes.data, capacity, ctypes.byref(n)), "hz_serve_poll")
    return [int(x) for x in ids[: n.value]]
