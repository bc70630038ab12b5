#!/usr/bin/env bash
# Builds the CPU oracle (test infrastructure) into oracle/liboracle.so.
# -ffp-contract=off keeps the float operation order of the restated kernels; -mfma lets the
# explicit std::fma of the texture lerps compile to the instruction.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
g++ -O2 -std=c++17 -mfma -ffp-contract=off -fno-fast-math -fopenmp -fPIC -shared \
    -o "$here/liboracle.so.tmp" "$here/oracle.cpp" "$here/oracle_capi.cpp"
mv -f "$here/liboracle.so.tmp" "$here/liboracle.so"
echo "built $here/liboracle.so"
