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
    "complete_{}", "clear_flags_{}", "set_device_{}",
]


def declared_symbols():
    text = HEADER.read_text()
    # Drop comments and the body of the per-program declaration macro.
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"#define HZ_DECLARE_PROGRAM_ABI\(suffix\)(?:.*\\\n)*.*\n", "", text)
    names = set(re.findall(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M))
    for suffix in re.findall(r"^HZ_DECLARE_PROGRAM_ABI\((\w+)\)", text, flags=re.M):
        names.update(p.format(suffix) for p in PROGRAM_ENTRY_POINTS)
    return sorted(names)


def test_header_parses_to_expected_surface():
    names = declared_symbols()
    for must in ("hz_last_error", "hz_collatz_run", "harmonize_alloc_device_bytes", "exec_program_collatz_async",
                 "alloc_program_collatz_event"):
        assert must in names
    assert len(names) >= 30


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
