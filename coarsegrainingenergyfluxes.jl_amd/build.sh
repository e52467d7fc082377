#!/usr/bin/env bash
# Builds libcgef_hip.so in-tree for gfx950. Packed FP32 instructions are disabled for the device
# code; host-side footprint construction is compiled without FP contraction so its tables match
# the reference's plain IEEE arithmetic.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
SRC="$HERE/csrc"
OUT="$HERE/cgef/libcgef_hip.so"
OBJ="$HERE/_build"
mkdir -p "$OBJ"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
ARCH="--offload-arch=gfx950"
DEVFLAGS="-Xclang -target-feature -Xclang -packed-fp32-ops"
COMMON="-O3 -fPIC -std=c++17 -I$HERE/../include"
"$HIPCC" $ARCH $COMMON $DEVFLAGS -c "$SRC/cgef_kernels.hip" -o "$OBJ/cgef_kernels.o"
"$HIPCC" $ARCH $COMMON $DEVFLAGS -c "$SRC/cgef_ring.hip" -o "$OBJ/cgef_ring.o"
"$HIPCC" $ARCH $COMMON $DEVFLAGS -c "$SRC/cgef_prefix.hip" -o "$OBJ/cgef_prefix.o"
"$HIPCC" $ARCH $COMMON $DEVFLAGS -c "$SRC/cgef_closure.hip" -o "$OBJ/cgef_closure.o"
"$HIPCC" $ARCH $COMMON $DEVFLAGS -c "$SRC/cgef_noalloc.hip" -o "$OBJ/cgef_noalloc.o"
"$HIPCC" $ARCH $COMMON -ffp-contract=off -c "$SRC/cgef_prefix_host.cpp" -o "$OBJ/cgef_prefix_host.o"
"$HIPCC" $ARCH $COMMON -ffp-contract=off -c "$SRC/cgef_ring_host.cpp" -o "$OBJ/cgef_ring_host.o"
"$HIPCC" $ARCH $COMMON -ffp-contract=off -c "$SRC/cgef_footprint.cpp" -o "$OBJ/cgef_footprint.o"
"$HIPCC" $ARCH $COMMON -ffp-contract=off -c "$SRC/cgef_api.cpp" -o "$OBJ/cgef_api.o"
"$HIPCC" $ARCH -shared -fPIC -o "$OUT.tmp" "$OBJ/cgef_kernels.o" "$OBJ/cgef_ring.o" "$OBJ/cgef_prefix.o" "$OBJ/cgef_closure.o" "$OBJ/cgef_noalloc.o" "$OBJ/cgef_prefix_host.o" "$OBJ/cgef_ring_host.o" "$OBJ/cgef_footprint.o" "$OBJ/cgef_api.o"
mv -f "$OUT.tmp" "$OUT"
echo "built $OUT"
