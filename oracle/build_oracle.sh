#!/usr/bin/env bash
# Build the oracle's C restatement into oracle/_build/liborc.so (git-ignored;
# it travels with the tree).  -O2, no fast-math, no FMA contraction, so the
# forward rounds like the reference's own x86-64 build.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
mkdir -p "$HERE/_build"
${CC:-gcc} -std=c99 -O2 -fPIC -fopenmp -ffp-contract=off -shared \
    -o "$HERE/_build/liborc.so.tmp" "$HERE/sf_oracle.c" -lm
mv -f "$HERE/_build/liborc.so.tmp" "$HERE/_build/liborc.so"
echo "built $HERE/_build/liborc.so"
