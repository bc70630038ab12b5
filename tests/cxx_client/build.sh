#!/bin/bash
# Builds the C++ adapter client (host code only: g++ against the HIP runtime and libartia.so).
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
root="$(cd "$here/../.." && pwd)"
rocm="${ROCM_PATH:-/opt/rocm}"
g++ -std=c++17 -O2 -Wall -Werror -D__HIP_PLATFORM_AMD__ -I "$root/include" -I "$rocm/include" \
    -o "$here/artia_client" "$here/client.cpp" \
    -L "$root/artiatomi_amd" -lartia -L "$rocm/lib" -lamdhip64 \
    -Wl,-rpath,'$ORIGIN/../../artiatomi_amd' -Wl,-rpath,"$rocm/lib"
echo "built $here/artia_client"
