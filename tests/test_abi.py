"""The C-ABI library loads and exports every entry point that
include/harmonize_amd.h declares (no GPU needed: nothing is called)."""

import ctypes
import re
from pathlib import Path

import pytest

import harmonize_amd as hz

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "harmonize_amd.h"

PROGRAM_ENTRY_POINTS = [
    "alloc_state_{}", "free_state_{}", "load_state_device_{}", "store_state_device_{}",
    "alloc_program_{}", "free_program_{}", "init_program_{}", "exec_program_{}",
    "complete_{}", "clear_flags_{}", "set_device_{}", "init_program_on_{}", "exec_program_on_{}",
]


STATE_LABEL_ENTRY_POINTS = [
    "store_pointer_state_device_{}", "load_state_group_{}", "store_state_group_{}", "store_pointer_state_group_{}",
    "load_state_thread_{}", "store_state_thread_{}", "store_pointer_state_thread_{}",
]
FIELD_ENTRY_POINTS = ["load_state_device_{f}_{s}", "store_state_device_{f}_{s}", "store_pointer_state_device_{f}_{s}"]


def state_field_lists(text):
    """{list macro name: [(field, is_pointer)]} from the header's X-macro field lists."""
    lists = {}
    for name, body in re.findall(r"#define (HZ_\w+_STATE_FIELDS)\(X, suffix\)((?:.*\\\n)*.*)\n", text):
        lists[name] = [(f, p == "1") for f, p in re.findall(r"X\(suffix, (\w+), ([01])\)", body)]
    return lists


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    lists = state_field_lists(text)
    # Drop macro definitions (their bodies are expanded below).
    text = re.sub(r"#define \w+\([^)]*\)(?:.*\\\n)*.*\n", "", text)
    names = set(re.findall(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M))
    for suffix in re.findall(r"^HZ_DECLARE_PROGRAM_ABI\((\w+)\)", text, flags=re.M):
        names.update(p.format(suffix) for p in PROGRAM_ENTRY_POINTS)
    for suffix, fields in re.findall(r"^HZ_DECLARE_STATE_LABELS\((\w+), (\w+)\)", text, flags=re.M):
        names.update(p.format(suffix) for p in STATE_LABEL_ENTRY_POINTS)
        for field, _ in lists[fields]:
            names.update(p.format(f=field, s=suffix) for p in FIELD_ENTRY_POINTS)
    return sorted(names)


def test_state_field_lists_match_the_c_structs():
    # Every listed field exists in its struct, in declaration order, and the
    # pointer flags agree with the declared types.
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    lists = state_field_lists(text)
    structs = {"HZ_COLLATZ_STATE_FIELDS": "hz_collatz_state", "HZ_PHOTON_STATE_FIELDS": "hz_photon_state",
               "HZ_BFS_STATE_FIELDS": "hz_bfs_state", "HZ_BARRIER_STATE_FIELDS": "hz_barrier_state"}
    assert set(lists) == set(structs)
    for macro, struct in structs.items():
        body = re.search(r"typedef struct %s \{(.*?)\}" % struct, text, flags=re.S).group(1)
        decls = []
        for line in body.split(";"):
            line = line.strip()
            if not line:
                continue
            ptr = "*" in line
            for name in re.findall(r"\*?\s*(\w+)\s*(?:,|$)", line.split(None, 1)[1] if " " in line else line):
                if name != "pad":
                    decls.append((name, ptr))
        assert lists[macro] == decls, macro


def test_header_parses_to_expected_surface():
    names = declared_symbols()
    for must in ("hz_last_error", "hz_collatz_run", "harmonize_alloc_device_bytes", "exec_program_collatz_async",
                 "alloc_program_collatz_event"):
        assert must in names
    assert len(names) >= 30
    for must in ("load_state_device_starts_collatz_async", "store_pointer_state_device_tally_photon_event",
                 "store_state_device_split_bfs_async", "load_state_group_barrier_async", "wall_clock_rate"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not hz.LIB_PATH.exists():
        pytest.fail(f"{hz.LIB_PATH} is not built")
    handle = ctypes.CDLL(str(hz.LIB_PATH))
    missing = [n for n in declared_symbols() if not hasattr(handle, n)]
    assert not missing, missing


def test_python_binding_signatures_cover_header():
    # Every name the Python binding types must exist in the header.
    declared = set(declared_symbols())
    assert set(hz._SIGNATURES) <= declared


def test_missing_library_fails_loudly():
    # No CPU fallback: without the built runtime every entry point raises.
    import subprocess
    import sys

    code = (
        "import harmonize_amd as hz, pathlib\n"
        "hz.LIB_PATH = pathlib.Path('/nonexistent/libharmonize_amd.so')\n"
        "hz._lib = None\n"
        "try:\n"
        "    hz.collatz_make_starts(1, 40, 16)\n"
        "except hz.HarmonizeError as e:\n"
        "    print('raised', 'not built' in str(e))\n"
    )
    proc = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                          cwd=str(ROOT))
    assert proc.stdout.strip() == "raised True", proc.stdout + proc.stderr


class _BfsState(ctypes.Structure):
    _fields_ = [("jobs", ctypes.c_void_p), ("job_iter", ctypes.c_void_p), ("job_count", ctypes.c_uint32),
                ("split", ctypes.c_uint32)]


@pytest.mark.parametrize("split", [0, (1 << 24) + 1, 0xFFFFFFFF])
def test_bfs_split_out_of_range_is_refused_before_any_launch(split):
    # A split threshold of 0 would cut a one-edge range forever and one above
    # 2^24 would overflow the wavefront's 32-bit prefix of range lengths; both
    # entry points refuse them before touching the device (no GPU here).
    lib = hz.lib()
    state = _BfsState(None, None, 1, split)
    h = ctypes.c_void_p()
    rc = lib.hz_instance_create(hz.KINDS["async"], hz.SPEC_BFS, 1024, ctypes.byref(state), ctypes.sizeof(state),
                                ctypes.byref(h))
    assert rc != 0 and h.value is None
    assert "split must be in [1, 2^24]" in hz.last_error()
    if split > (1 << 24):
        cfg = hz.RunConfig()
        cfg.spec_param = split
        rc = lib.hz_bfs_batch_run(None, hz.KINDS["async"], ctypes.byref(cfg), None)
        assert rc != 0
        assert "split threshold" in hz.last_error()


@pytest.mark.parametrize("field,value", [("arena", 16385), ("queues", 17)])
def test_noalloc_refuses_sizes_beyond_its_static_array_before_any_launch(field, value):
    # The allocation-free entry point's buffers live in a fixed device array:
    # configurations that would not fit are refused before the device is
    # touched (no GPU here).
    cfg = hz.RunConfig()
    setattr(cfg, field, value)
    st = hz.RunStats()
    rc = hz.lib().hz_collatz_run_noalloc(None, 0, None, ctypes.byref(cfg), ctypes.byref(st))
    assert rc == -1
    assert "static maximum" in hz.last_error()


def test_audit_catches_planted_inconsistencies():
    # The HZ_RUN_AUDIT check accepts a consistent snapshot of queues and arena
    # and names each planted fault (host code only, no GPU).
    buf = ctypes.create_string_buffer(4096)
    missed = hz.lib().hz_audit_selftest(buf, len(buf))
    assert missed == 0, buf.value.decode()


def test_tune_cache_round_trip(tmp_path):
    # The tuning cache file (host only): entries round-trip, a second store
    # under the same key replaces the first, other keys are untouched, and a
    # disabled cache finds nothing.
    path = tmp_path / "tune.txt"
    hz.tune_cache_set_path(str(path))
    try:
        assert hz.tune_cache_lookup("bfs_batch_async|gfx950|256|graphs=64") is None
        a = dict(groups=2048, arena=466741, queues=16, coalesce=32, block_groups=2, flags=1, spec_param=128, stash=8)
        hz.tune_cache_store("bfs_batch_async|gfx950|256|graphs=64", a, 37.6)
        hz.tune_cache_store("collatz_async|gfx950|256|starts=2^26", dict(queues=16), 67.0)
        b = dict(a, spec_param=64)
        hz.tune_cache_store("bfs_batch_async|gfx950|256|graphs=64", b, 36.9)
        got, ms = hz.tune_cache_lookup("bfs_batch_async|gfx950|256|graphs=64")
        assert ms == 36.9 and {k: v for k, v in got.items() if v} == b
        got, ms = hz.tune_cache_lookup("collatz_async|gfx950|256|starts=2^26")
        assert got["queues"] == 16 and ms == 67.0
        assert len(path.read_text().splitlines()) == 2
        with pytest.raises(hz.HarmonizeError):
            hz.tune_cache_store("bad\tkey", a, 1.0)
        hz.tune_cache_set_path(None)
        assert hz.tune_cache_lookup("collatz_async|gfx950|256|starts=2^26") is None
    finally:
        hz.tune_cache_set_path(None)
