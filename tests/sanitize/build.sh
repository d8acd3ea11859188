#!/bin/bash
# Builds the library's C ABI with AddressSanitizer and UndefinedBehaviorSanitizer on the host code only (the
# device code is compiled as usual for gfx950) and links the host check program tests/sanitize/abi_errors.c.
# Usage: build.sh OUTDIR. Runs on a machine without a GPU; the program makes no device calls.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
OUT="$1"
mkdir -p "$OUT"
CSRC="$ROOT/nonlinear-fiber-homogenization_amd/csrc"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
SAN="-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-omit-frame-pointer"
FLAGS="--offload-arch=gfx950 -O1 -g -std=c++17 -fPIC -Xclang -target-feature -Xclang -packed-fp32-ops -Wno-unused-result"
pids=()
for src in nlfh_matvec nlfh_fft nlfh_fft_fast nlfh_blas nlfh_capi nlfh_net nlfh_density nlfh_finite; do
  $HIPCC $FLAGS $SAN -I "$ROOT/include" -I "$CSRC" -c "$CSRC/$src.hip" -o "$OUT/$src.o" 2> "$OUT/$src.log" &
  pids+=($!)
done
for p in "${pids[@]}"; do wait "$p"; done
$HIPCC $FLAGS $SAN -I "$ROOT/include" -x c++ -c "$HERE/abi_errors.c" -o "$OUT/abi_errors.o" 2> "$OUT/abi_errors.log"
$HIPCC --offload-arch=gfx950 $SAN "$OUT"/*.o -o "$OUT/abi_errors"
