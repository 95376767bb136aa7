#!/usr/bin/env bash
# Builds lib/libreactiveTransport.so (HIP, gfx950) from csrc/.
# Rounding policy: IEEE FP64, no contraction (-ffp-contract=off), no fast-math;
# search.hip (the FP32 candidate search only) adds -freciprocal-math -fapprox-func.
set -euo pipefail
here="$(cd -- "$(dirname -- "${BASH_SOURCE[0]}")" && pwd)"
hipcc="${HIPCC:-/opt/rocm/bin/hipcc}"
mkdir -p "$here/lib" "$here/_build"
flags=(--offload-arch=gfx950 -O3 -std=c++17 -fPIC -ffp-contract=off -fno-fast-math
       -Xclang -target-feature -Xclang -packed-fp32-ops
       -I"$here/../include" -Wall -Wno-unused-result)
objs=()
pids=()
for src in "$here"/csrc/*.hip; do
  obj="$here/_build/$(basename "${src%.hip}").o"
  rm -f "$obj"
  extra=()
  # the FP32 candidate search may use approximate reciprocals and transcendentals
  if [ "$(basename "$src")" = "search.hip" ]; then extra=(-freciprocal-math -fapprox-func); fi
  "$hipcc" "${flags[@]}" "${extra[@]}" -c "$src" -o "$obj" 2> >(grep -v "packed-fp32-ops" >&2) &
  objs+=("$obj")
  pids+=("$!")
done
failed=0
for pid in "${pids[@]}"; do
  wait "$pid" || failed=1
done
if [ "$failed" -ne 0 ]; then
  echo "build.sh: a HIP compile failed" >&2
  exit 1
fi
"$hipcc" --offload-arch=gfx950 -shared -fPIC "${objs[@]}" -o "$here/lib/libreactiveTransport.so.tmp"
mv "$here/lib/libreactiveTransport.so.tmp" "$here/lib/libreactiveTransport.so"
