"""Builds libharmonize_amd.so (gfx950 HIP) in-tree, and the CPU oracle.

    python -m harmonize_amd.build            # runtime library + oracle
    python -m harmonize_amd.build --runtime  # runtime library only

Each csrc/*.hip file is compiled to an object in _build/ (in parallel) and
linked into harmonize_amd/libharmonize_amd.so. The oracle (oracle/*.cpp,
test infrastructure) is compiled with g++ into oracle/liboracle.so.
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
BUILD = ROOT / "_build"
LIB = PKG / "libharmonize_amd.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"

ARCH = os.environ.get("HZ_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-std=c++17",
    "-O3",
    f"--offload-arch={ARCH}",
    "-fPIC",
    # No packed FP32 instructions: some of their forms have given wrong
    # results on MI355X without any error.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Host and device evaluate specification arithmetic identically.
    "-ffp-contract=off",
    "-munsafe-fp-atomics",
    "-Wno-unused-command-line-argument",
    f"-I{PKG / 'include'}",
    f"-I{PKG / 'specs'}",
    f"-I{ROOT / 'include'}",
]

HOST_FLAGS = [
    "-std=c++17",
    "-O2",
    "-fopenmp",
    "-fPIC",
    "-ffp-contract=off",
    f"-I{PKG / 'include'}",
    f"-I{PKG / 'specs'}",
    f"-I{ORACLE_DIR}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the gfx950 runtime cannot be built")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line and line.strip()
    )
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def build_runtime(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted((PKG / "csrc").glob("*.hip"))
    if not sources:
        raise RuntimeError("no HIP sources found")
    cc = hipcc()
    objects = [BUILD / (s.stem + ".o") for s in sources]
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(sources)))) as pool:
        futures = [pool.submit(_run, [cc, *HIP_FLAGS, "-c", str(s), "-o", str(o)]) for s, o in zip(sources, objects)]
        for f in futures:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([cc, "-shared", f"--offload-arch={ARCH}", "-fPIC", *map(str, objects), "-o", str(tmp)])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    sources = sorted(ORACLE_DIR.glob("*.cpp"))
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run(["g++", *HOST_FLAGS, "-shared", *map(str, sources), "-o", str(tmp)])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


REFERENCE = Path(os.environ.get("HZ_REFERENCE", "/root/reference"))


def build_reference_examples() -> Path | None:
    """Builds the reference's own Collatz examples against this runtime's
    headers into oracle/_ref/ (oracle/build_ref.sh) when the reference
    sources are present; returns the output folder, or None."""
    if not (REFERENCE / "examples" / "cpp" / "collaz" / "collaz_hrm.cpp").exists():
        return None
    _run(["bash", str(ORACLE_DIR / "build_ref.sh")])
    return ORACLE_DIR / "_ref"


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--runtime", action="store_true", help="build only the runtime library")
    ap.add_argument("--oracle", action="store_true", help="build only the oracle")
    ap.add_argument("-j", "--jobs", type=int, default=8)
    args = ap.parse_args(argv)
    both = not (args.runtime or args.oracle)
    if args.runtime or both:
        print("built", build_runtime(args.jobs))
    if args.oracle or both:
        print("built", build_oracle())
        ref = build_reference_examples()
        if ref:
            print("built", ref)
    return 0


if __name__ == "__main__":
    sys.exit(main())
