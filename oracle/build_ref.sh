#!/usr/bin/env bash
# Compile the reference's own SF forward, parameter-file reader and neighbour
# list from the sources where they lie under /root/reference, together with
# oracle/ref_driver.cpp, into oracle/_ref/libsfref.so.  Nothing is copied out
# of the reference; outputs go to oracle/_ref/ only (git-ignored).
#
# The reference objects get -fvisibility=hidden and the link uses
# --gc-sections: the Enzyme gradient entry points (which need the Enzyme
# plugin, absent here) and the Bispectrum/SOAP/Xi arms are never referenced
# from the driver and are dropped, so the link closes without stand-ins.
# --no-undefined checks that.
set -euo pipefail
REF=${REF:-/root/reference}
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/_ref"
if [ ! -d "$REF/src" ]; then
    echo "build_ref.sh: $REF not present; skipping the reference build" >&2
    exit 0
fi
mkdir -p "$OUT"
CXX=${CXX:-g++}
FLAGS="-std=c++17 -O2 -fPIC -ffunction-sections -fdata-sections -fvisibility=hidden -w"
$CXX $FLAGS -I"$REF/include" -c "$REF/src/SymmetryFunctions.cpp" -o "$OUT/SymmetryFunctions.o"
$CXX $FLAGS -I"$REF/include" -c "$REF/src/Descriptors.cpp" -o "$OUT/Descriptors.o"
$CXX $FLAGS -I"$REF/include" -c "$REF/src/helper.cpp" -o "$OUT/helper.o"
$CXX $FLAGS -I"$REF/python_bindings/neighbor" -c "$REF/python_bindings/neighbor/neighbor_list.cpp" \
    -o "$OUT/neighbor_list.o"
$CXX $FLAGS -I"$REF/include" -I"$REF/python_bindings/neighbor" -c "$HERE/ref_driver.cpp" \
    -o "$OUT/ref_driver.o"
$CXX -shared -o "$OUT/libsfref.so.tmp" "$OUT/ref_driver.o" "$OUT/SymmetryFunctions.o" \
    "$OUT/Descriptors.o" "$OUT/helper.o" "$OUT/neighbor_list.o" \
    -Wl,--gc-sections -Wl,--no-undefined
mv -f "$OUT/libsfref.so.tmp" "$OUT/libsfref.so"
rm -f "$OUT"/*.o
echo "built $OUT/libsfref.so"
