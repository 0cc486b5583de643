#!/bin/bash
# Builds libelph_vertex.so (gfx950) in-tree.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/../libelph_vertex.so"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
"$HIPCC" --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -Wall -Wno-unused-result -pthread \
  -o "$OUT.tmp" "$HERE/elph_vertex.hip" "$HERE/serve.hip" "$HERE/transport.hip" "$HERE/host_setup.cpp" "$HERE/host_state.cpp" "$HERE/g2_writer.cpp"
mv "$OUT.tmp" "$OUT"
echo "built $OUT"
