#!/bin/bash
# Builds the C oracle (test infrastructure) into oracle/_build/liboracle.so and,
# when the reference tree is present, the reference parser dump in oracle/_ref/.
# Serial FP64, no FMA contraction, no -march (the same binary runs on the GPU
# box's host CPU).
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$HERE/_build"
gcc -std=c11 -O3 -ffp-contract=off -fPIC -shared -Wall -Wno-unused-parameter -Wno-unknown-pragmas \
    -o "$HERE/_build/liboracle.so.tmp" "$HERE/orc_chem.c" "$HERE/orc_flow.c" -lm
mv -f "$HERE/_build/liboracle.so.tmp" "$HERE/_build/liboracle.so"
bash "$HERE/build_ref.sh"
