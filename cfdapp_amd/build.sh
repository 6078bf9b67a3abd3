#!/bin/sh
# Builds libcfdhip.so (HIP for gfx950) in-tree next to this script: every
# .hip translation unit compiled in parallel into _build/, then linked.
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
OBJ="$HERE/_build"
JOBS="${BUILD_JOBS:-8}"
mkdir -p "$OBJ"
export HIPCC OBJ
ls "$HERE"/csrc/*.hip | xargs -P "$JOBS" -I{} sh -c '
  f="{}"; o="$OBJ/$(basename "$f" .hip).o"
  "$HIPCC" -std=c++20 -O3 --offload-arch=gfx950 -fPIC -c \
    -ffp-contract=off -Xclang -target-feature -Xclang -packed-fp32-ops \
    -Wall -Wno-unused-function "$f" -o "$o"'
"$HIPCC" -shared --offload-arch=gfx950 "$OBJ"/*.o -o "$HERE/libcfdhip.so.tmp"
mv "$HERE/libcfdhip.so.tmp" "$HERE/libcfdhip.so"
# C++20 host-adapter test driver (include/cfdhip.hpp over the C ABI)
g++ -std=c++20 -O2 -Wall -Wextra -I"$HERE/../include" "$HERE/../tests/cpp/test_host_adapter.cpp" \
  -L"$HERE" -lcfdhip -Wl,-rpath,'$ORIGIN/../../cfdapp_amd' -o "$HERE/../tests/cpp/test_host_adapter"
