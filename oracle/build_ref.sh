#!/bin/bash
# Compiles the one piece of the reference's hot path that builds from its own
# sources with the host compiler: ParameterOrganizer (knot packing, used by
# mexgateway/task to move knots in and out of the solver's layout).  Sources are
# read where they lie under /root/reference; the output goes to oracle/_ref/.
# Everything else on the path needs CUDA, Eigen or MATLAB headers and is
# unbuildable here.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
SRC=/root/reference/setup/source
if [ ! -d "$SRC" ]; then
  echo "reference sources not present; skipping oracle/_ref build" >&2
  exit 0
fi
mkdir -p "$HERE/_ref"
g++ -std=c++11 -O2 -fPIC -shared -D_DEBUG -w \
  -I"$SRC" \
  "$SRC/interface/parameterorganizer/parameterorganizer.cpp" \
  "$SRC/helper/memtracker/memtracker.cpp" \
  "$HERE/ref_paramorg_shim.cpp" \
  -o "$HERE/_ref/libref_paramorg.so"
echo "built $HERE/_ref/libref_paramorg.so"
