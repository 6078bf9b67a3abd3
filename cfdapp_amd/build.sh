#!/bin/sh
# Builds libcfdhip.so (HIP for gfx950) in-tree next to this script.
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
"$HIPCC" -std=c++20 -O3 --offload-arch=gfx950 -fPIC -shared \
  -ffp-contract=off -Xclang -target-feature -Xclang -packed-fp32-ops \
  -Wall -Wno-unused-function \
  "$HERE"/csrc/*.hip -o "$HERE/libcfdhip.so.tmp"
mv "$HERE/libcfdhip.so.tmp" "$HERE/libcfdhip.so"
# C++20 host-adapter test driver (include/cfdhip.hpp over the C ABI)
g++ -std=c++20 -O2 -Wall -Wextra -I"$HERE/../include" "$HERE/../tests/cpp/test_host_adapter.cpp" \
  -L"$HERE" -lcfdhip -Wl,-rpath,'$ORIGIN/../../cfdapp_amd' -o "$HERE/../tests/cpp/test_host_adapter"
