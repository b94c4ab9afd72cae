"""The reference's own example programs (examples/cpp/collaz) built unchanged
against this runtime's harmonize.h by oracle/build_ref.sh: on the GPU they
run to completion and their built-in checker (collaz_hrm.cpp:157-169) finds
no bad output, for the EventProgram as written and for the AsyncProgram
refactor. Skipped where the binaries were not built (no reference sources)."""

import subprocess
from pathlib import Path

import pytest

REF = Path(__file__).resolve().parents[1] / "oracle" / "_ref"


def run(binary, *args, timeout=120):
    exe = REF / binary
    if not exe.exists():
        pytest.skip(f"{exe} not built (reference sources absent when building)")
    proc = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=timeout)
    return proc


@pytest.mark.gpu
@pytest.mark.parametrize("binary,wg", [("clz_hrm_event", 2048), ("clz_hrm_async", 1024)])
def test_reference_collatz_program_runs_unchanged(binary, wg):
    proc = run(binary, "-wg_count", wg, "-limit", 1 << 20)
    assert proc.returncode == 0, proc.stderr[-2000:]
    assert "Bad output" not in proc.stdout
    assert "Runtime:" in proc.stdout


@pytest.mark.gpu
def test_reference_plain_collatz_kernel_runs():
    # limit a multiple of the workgroup size: the example's kernel guards only
    # each workgroup's first index (collaz_std.cpp:6-9).
    proc = run("clz_std", "-wg_count", 4096, "-wg_size", 64, "-limit", 1 << 20)
    assert proc.returncode == 0, proc.stderr[-2000:]
    assert "Bad output" not in proc.stdout
    assert "Runtime:" in proc.stdout
