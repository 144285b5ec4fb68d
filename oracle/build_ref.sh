#!/bin/bash
# Builds the reference's own brick/region test programs and the brick trace
# driver from the reference sources where they lie (header-only C++), into
# oracle/_ref/.  Needs /root/reference (absent on GPU machines: nothing there
# uses these binaries).
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
src=/root/reference/FluxFX_Source/native
out="$here/_ref"
mkdir -p "$out"
g++ -std=c++17 -O1 -I"$src" "$src/bricks_test.cpp" -o "$out/bricks_test"
g++ -std=c++17 -O1 -I"$src" "$src/regions_test.cpp" -o "$out/regions_test"
g++ -std=c++17 -O1 -I"$src" "$here/ref/brick_trace.cpp" -o "$out/brick_trace"
