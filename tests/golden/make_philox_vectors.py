"""Writes philox4x32_10.json: Philox4x32-10 outputs from rocRAND's open-source
implementation (rocrand_philox4x32_10.h, shipped with ROCm), an implementation
independent of the one in harmonize_amd/specs/philox.h.

The vectors are the three known-answer tests of the Random123 distribution
(counter/key all zeros, all ones, and the digits of pi) plus 61 seeded
random counter/key pairs. Run from the repository root:

    python tests/golden/make_philox_vectors.py
"""

import json
import os
import random
import subprocess
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent

PROGRAM = r"""
#include <cstdio>
#include <cstdlib>
#include <rocrand/rocrand_philox4x32_10.h>

// rocRAND keeps its ten-round function protected; expose it.
struct Probe : rocrand_device::philox4x32_10_engine {
    using rocrand_device::philox4x32_10_engine::ten_rounds;
};

int main(int argc, char** argv) {
    unsigned c0, c1, c2, c3, k0, k1;
    while (std::scanf("%x %x %x %x %x %x", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        Probe p;
        uint4 r = p.ten_rounds(uint4{c0, c1, c2, c3}, uint2{k0, k1});
        std::printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
"""


def main():
    rng = random.Random(20261202)
    cases = [
        ([0, 0, 0, 0], [0, 0]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]),
    ]
    for _ in range(61):
        cases.append(([rng.getrandbits(32) for _ in range(4)], [rng.getrandbits(32) for _ in range(2)]))
    with tempfile.TemporaryDirectory() as tmp:
        src = Path(tmp) / "probe.cpp"
        exe = Path(tmp) / "probe"
        src.write_text(PROGRAM)
        subprocess.run(["hipcc", "-O1", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", str(src), "-o", str(exe)],
                       check=True)
        stdin = "".join(" ".join(f"{x:08x}" for x in c + k) + "\n" for c, k in cases)
        out = subprocess.run([str(exe)], input=stdin, capture_output=True, text=True, check=True).stdout.split("\n")
    vectors = []
    for (c, k), line in zip(cases, out):
        vectors.append({"counter": c, "key": k, "out": [int(x, 16) for x in line.split()]})
    (HERE / "philox4x32_10.json").write_text(json.dumps({"source": "rocrand_philox4x32_10.h ten_rounds",
                                                         "vectors": vectors}, indent=1))
    print("wrote", len(vectors), "vectors")


if __name__ == "__main__":
    os.chdir(HERE.parent.parent)
    main()
