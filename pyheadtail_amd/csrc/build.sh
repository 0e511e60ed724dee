#!/bin/bash
# Builds libphx.so (gfx950) next to the package from the sources in this
# directory.  Usage: build.sh [output_dir]
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
out="${1:-$here/..}"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
ARCH="${PHX_ARCH:-gfx950}"
FLAGS=(-O3 -std=c++17 -fPIC --offload-arch="$ARCH" -ffp-contract=off
       -Xclang -target-feature -Xclang -packed-fp32-ops
       -Wall -Wno-unused-function -Wno-unused-variable)
# PHX_DEFINES: extra -D options (alternative compile-time choices);
# PHX_BUILD_DIR: where the objects go
read -r -a defs <<< "${PHX_DEFINES:-}"
bdir="${PHX_BUILD_DIR:-$here/_build}"
objs=()
pids=()
mkdir -p "$bdir"
for src in phx_api phx_stats phx_track phx_wake phx_engine phx_batch phx_spacecharge phx_sort; do
  "$HIPCC" "${FLAGS[@]}" "${defs[@]}" -c "$here/$src.hip" -o "$bdir/$src.o" &
  pids+=($!)
  objs+=("$bdir/$src.o")
done
for p in "${pids[@]}"; do wait "$p"; done
"$HIPCC" -shared --offload-arch="$ARCH" -o "$out/libphx.so.tmp" "${objs[@]}"
mv "$out/libphx.so.tmp" "$out/libphx.so"
echo "built $out/libphx.so"
