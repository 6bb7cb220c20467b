#!/bin/bash
# Builds, in-tree:
#   libommhip.so  HIP kernels + C-ABI (hipcc, gfx950)
#   libommf.so    Fortran MatrixSwitch / omm() drop-in over the C-ABI (amdflang)
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(cd "$HERE/.." && pwd)"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
FC="${FC_OMM:-/opt/rocm/bin/amdflang}"
OBJDIR="$HERE/_build"
mkdir -p "$OBJDIR/f"
SRC=("$HERE"/csrc/api.hip "$HERE"/csrc/gemm.hip "$HERE"/csrc/spmm.hip "$HERE"/csrc/elementwise.hip "$HERE"/csrc/adjoint.hip "$HERE"/csrc/zgemm.hip "$HERE"/csrc/dense_la.hip "$HERE"/csrc/batch.hip "$HERE"/csrc/lattice.cpp)
FLAGS=(--offload-arch=gfx950 -O3 -fPIC -std=c++17 -I"$ROOT/include" -I"$HERE/csrc"
       -Xclang -target-feature -Xclang -packed-fp32-ops)
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
"$HIPCC" --offload-arch=gfx950 -shared -fPIC -o "$HERE/libommhip.so.tmp" "${objs[@]}"
# A kernel whose host launch stub the compiler failed to emit leaves an undefined
# __device_stub__ symbol that only shows when the library is loaded: refuse such a build.
if nm -u "$HERE/libommhip.so.tmp" | grep -q "__device_stub__"; then
  echo "build.sh: undefined kernel launch stubs in libommhip.so:" >&2
  nm -u -C "$HERE/libommhip.so.tmp" | grep "__device_stub__" >&2
  rm -f "$HERE/libommhip.so.tmp"
  exit 1
fi
mv "$HERE/libommhip.so.tmp" "$HERE/libommhip.so"
echo "built $HERE/libommhip.so"

# Fortran drop-in (module order follows the USE graph)
FOBJS=()
for f in omx_capi MatrixSwitch omm_hip_ops omm omm_c; do
  "$FC" -cpp -O2 -fPIC -J"$OBJDIR/f" -c "$HERE/fortran/$f.F90" -o "$OBJDIR/f/$f.o"
  FOBJS+=("$OBJDIR/f/$f.o")
done
"$FC" -shared -fPIC -o "$HERE/libommf.so.tmp" "${FOBJS[@]}" -L"$HERE" -lommhip -Wl,-rpath,'$ORIGIN'
mv "$HERE/libommf.so.tmp" "$HERE/libommf.so"
echo "built $HERE/libommf.so"

# Fortran example programs (link against the drop-in)
for ex in example_gamma_hip example_sdden_hip example_kpoints_hip; do
  "$FC" -cpp -O2 -I"$OBJDIR/f" -o "$HERE/examples/$ex" "$HERE/examples/$ex.F90" -L"$HERE" -lommf -lommhip \
    -Wl,-rpath,'$ORIGIN/..'
done
echo "built examples"
