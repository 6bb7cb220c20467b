#!/bin/bash
# Builds libommhip.so (HIP kernels + C-ABI) for gfx950 in-tree.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(cd "$HERE/.." && pwd)"
OUT="$HERE/libommhip.so"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
SRC=("$HERE"/csrc/api.hip "$HERE"/csrc/gemm.hip "$HERE"/csrc/spmm.hip "$HERE"/csrc/elementwise.hip "$HERE"/csrc/lattice.cpp)
OBJDIR="$HERE/_build"
mkdir -p "$OBJDIR"
FLAGS=(--offload-arch=gfx950 -O3 -fPIC -std=c++17 -I"$ROOT/include" -I"$HERE/csrc"
       -Xclang -target-feature -Xclang -packed-fp32-ops -Wno-unused-result)
pids=()
objs=()
for s in "${SRC[@]}"; do
  o="$OBJDIR/$(basename "$s").o"
  objs+=("$o")
  if [[ "$s" == *.cpp ]]; then
    g++ -O3 -fPIC -std=c++17 -I"$ROOT/include" -I"$HERE/csrc" -c "$s" -o "$o" &
  else
    "$HIPCC" "${FLAGS[@]}" -c "$s" -o "$o" 2> >(grep -v "packed-fp32-ops" >&2) &
  fi
  pids+=($!)
done
for p in "${pids[@]}"; do wait "$p"; done
"$HIPCC" --offload-arch=gfx950 -shared -fPIC -o "$OUT.tmp" "${objs[@]}"
mv "$OUT.tmp" "$OUT"
echo "built $OUT"
