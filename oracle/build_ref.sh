#!/bin/bash
# Builds the reference's own example programs, unchanged, against this
# repository's harmonize.h (the reference's Makefiles force-include its
# harmonize.h the same way, examples/cpp/collaz/Makefile,
# examples/cpp/example_neutron/Makefile). Outputs go to oracle/_ref/ only.
# The Collatz async build redefines EventProgram as AsyncProgram before
# including the example: the reference README's one-line refactor.
#   clz_hrm_event   examples/cpp/collaz/collaz_hrm.cpp as written (EventProgram)
#   clz_hrm_async   the same file with ProgType = AsyncProgram<MyProgramSpec>
#   clz_std         examples/cpp/collaz/collaz_std.cpp (one thread per start)
#   neut            examples/cpp/example_neutron/program.cpp as written; it
#                   instantiates both program types (EventProgram by default,
#                   AsyncProgram with -async)
set -euo pipefail
REF=${HZ_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
REPO=$(dirname "$HERE")
OUT="$HERE/_ref"
SRC="$REF/examples/cpp/collaz"
if [ ! -f "$SRC/collaz_hrm.cpp" ]; then
    echo "reference examples not found under $REF: nothing to build"
    exit 0
fi
mkdir -p "$OUT"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
FLAGS="-std=c++17 -x hip --offload-arch=gfx950 -O3 -Xclang -target-feature -Xclang -packed-fp32-ops -Wno-unused-command-line-argument -I$REPO/harmonize_amd/include"
$HIPCC $FLAGS -include "$REPO/harmonize_amd/include/harmonize.h" "$SRC/collaz_hrm.cpp" -o "$OUT/clz_hrm_event" 2>&1 | grep -v packed-fp32 || true
printf '#include "harmonize.h"\n#define EventProgram AsyncProgram\n#include "%s/collaz_hrm.cpp"\n' "$SRC" > "$OUT/clz_hrm_async.cpp"
$HIPCC $FLAGS "$OUT/clz_hrm_async.cpp" -o "$OUT/clz_hrm_async" 2>&1 | grep -v packed-fp32 || true
$HIPCC $FLAGS -include "$REPO/harmonize_amd/include/harmonize.h" "$SRC/collaz_std.cpp" -o "$OUT/clz_std" 2>&1 | grep -v packed-fp32 || true
NEUT="$REF/examples/cpp/example_neutron"
$HIPCC $FLAGS -include "$REPO/harmonize_amd/include/harmonize.h" "$NEUT/program.cpp" -o "$OUT/neut" 2>&1 | grep -v packed-fp32 || true
for b in clz_hrm_event clz_hrm_async clz_std neut; do
    [ -x "$OUT/$b" ] || { echo "failed to build $b"; exit 1; }
done
echo "built $OUT/clz_hrm_event $OUT/clz_hrm_async $OUT/clz_std $OUT/neut"
