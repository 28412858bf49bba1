#!/bin/sh
# Builds libpp_hip.so for gfx950 next to the Python package (paulistrings_amd/libpp_hip.so).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/../paulistrings_amd/libpp_hip.so"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
FLAGS="--offload-arch=gfx950 -O3 -std=c++17 -fPIC -ffp-contract=off -Xclang -target-feature -Xclang -packed-fp32-ops -Wall -Wno-unused-function"
$HIPCC $FLAGS -shared -o "$OUT.tmp" "$HERE/pp_kernels.hip" "$HERE/pp_topn.hip" "$HERE/pp_learned.hip" "$HERE/pp_membench.hip" "$HERE/pp_class.hip" "$HERE/pp_host.hip"
mv "$OUT.tmp" "$OUT"
