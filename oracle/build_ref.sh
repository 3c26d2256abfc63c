#!/usr/bin/env bash
# Build the reference checkers from the reference's own sources, where they lie.
#   oracle/_ref/librefpoint.so : the reference's inline EOS / flux / eigen / WENO kernels
#                                (include/Equations.hpp, include/Utility.hpp) behind C entry
#                                points (oracle/ref_point.cpp)
#   oracle/_ref/refsolver      : the reference's unmodified 4-rank solver (src/WENOFV.cpp)
#                                driven by oracle/ref_solver.cpp; needs the MPI found in
#                                this development container (/opt/conda)
#   oracle/_ref/refmain_gpu    : the reference's unmodified main program (src/main.cpp)
#                                built against this build's CWENOFV facade and libwenofv.so
# Flags follow the reference makefile (-O3, no -march: no FMA).  Outputs go only to
# oracle/_ref/ (git-ignored, and excluded from the tree uploaded to GPU machines by
# .gpurunignore).  Development container only: neither the reference nor anything built
# from it goes to the GPU machines, where only the committed fixtures are used.
set -euo pipefail
REF=${WENOFV_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
[ -d "$REF/include" ] || { echo "reference not present at $REF; skipping"; exit 0; }
mkdir -p "$OUT"
CXXFLAGS="-std=c++17 -O3 -w -I$REF/include -I$REF/third_party/scTools"
g++ $CXXFLAGS -fPIC -shared "$HERE/ref_point.cpp" -o "$OUT/librefpoint.so"
echo "built $OUT/librefpoint.so"
MPI_INC=${MPI_INC:-/opt/conda/include}
MPI_LIB=${MPI_LIB:-/opt/conda/lib}
if [ -f "$MPI_INC/mpi.h" ] && [ -f "$MPI_LIB/libmpi.so" ]; then
  g++ $CXXFLAGS -I"$MPI_INC" -c "$REF/src/WENOFV.cpp" -o "$OUT/WENOFV.o"
  g++ $CXXFLAGS -I"$MPI_INC" -c "$HERE/ref_solver.cpp" -o "$OUT/ref_solver.o"
  g++ -o "$OUT/refsolver" "$OUT/ref_solver.o" "$OUT/WENOFV.o" "$MPI_LIB/libmpi.so" \
      -Wl,-rpath,/usr/lib/x86_64-linux-gnu:"$MPI_LIB"
  echo "built $OUT/refsolver"
  # The reference's own main program (src/main.cpp, unmodified, compiled where it lies)
  # against this build's CWENOFV facade: its relative includes ("../include/WENOFV.hpp",
  # "../include/ReadConfig.hpp", "../third_party/...") resolve through a scratch tree of
  # symlinks to the facade headers and the reference's scTools; no source is copied.
  PKG="$(cd "$HERE/.." && pwd)/two-phase-3equation-2d-pal-column_amd"
  if [ -f "$PKG/libwenofv.so" ]; then
    T=$(mktemp -d)
    mkdir -p "$T/src" "$T/include"
    ln -s "$REF/src/main.cpp" "$T/src/main.cpp"
    ln -s "$REF/third_party" "$T/third_party"
    for h in WENOFV.hpp WENOFV_facade.hpp WENOFV_output.hpp ReadConfig.hpp Equations.hpp wenofv_array.h; do ln -s "$PKG/csrc/$h" "$T/include/$h"; done
    g++ -std=c++17 -O2 -w -I"$HERE/../include" -I"$MPI_INC" -c "$T/src/main.cpp" -o "$OUT/refmain.o"
    g++ -o "$OUT/refmain_gpu" "$OUT/refmain.o" -L"$PKG" -lwenofv "$MPI_LIB/libmpi.so" \
        -Wl,-rpath,'$ORIGIN/../../two-phase-3equation-2d-pal-column_amd':/usr/lib/x86_64-linux-gnu:"$MPI_LIB"
    rm -rf "$T"
    echo "built $OUT/refmain_gpu"
  fi
else
  echo "MPI not found; refsolver not built"
fi
