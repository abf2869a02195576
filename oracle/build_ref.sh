#!/bin/bash
# Builds oracle/_ref/ref_chem_dump: the reference's own host-side CHEMKIN /
# therm.dat / tran.dat / setup-file readers, compiled unmodified from their
# sources under /root/reference with g++, linked with this directory's
# ref_chem_dump.cpp driver.  Headers come from the reference tree (its vendored
# header-only fmt) and the image's MPICH (/opt/conda/include); the one MPI
# symbol the parsers reference on an unreachable abort path
# (cfd::MpiParallel::exit, defined in a CUDA translation unit) is left
# unresolved at link time.  Outputs go to oracle/_ref/ only.
set -euo pipefail
REF=${REF:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
if [ ! -d "$REF/src" ]; then
  echo "reference tree not present at $REF; skipping oracle/_ref build" >&2
  exit 0
fi
mkdir -p "$OUT"
FLAGS="-std=c++20 -O2 -DMAX_SPEC_NUMBER=20 -DMAX_REAC_NUMBER=50 -DFMT_HEADER_ONLY -I$REF/src -I$REF/depends/include -I/opt/conda/include"
objs=()
for f in ChemData.cpp Element.cpp Parameter.cpp gxl_lib/MyString.cpp; do
  o="$OUT/$(basename "$f" .cpp).o"
  g++ $FLAGS -w -c "$REF/src/$f" -o "$o"
  objs+=("$o")
done
g++ $FLAGS -c "$HERE/ref_chem_dump.cpp" -o "$OUT/ref_chem_dump.o"
g++ -o "$OUT/ref_chem_dump" "$OUT/ref_chem_dump.o" "${objs[@]}" \
  -L/opt/conda/lib -Wl,-rpath,/opt/conda/lib -lmpi -Wl,--unresolved-symbols=ignore-in-object-files
echo "built $OUT/ref_chem_dump"
