#!/bin/bash
# Compile the reference's own FFT and solver routines from their sources in the reference
# tree, plus the project test program oracle/refdrv.f90, into oracle/_ref/ (checker only).
# The missing FOU3D.f90 symbols and the MPI entry points stay unresolved: that program never
# reaches them, and the library is loaded lazily by tests/golden/make_golden.py.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${LH_REFERENCE:-/root/reference}"
OUT="$HERE/_ref"
FC="${FC:-/opt/rocm/bin/amdflang}"
MPIINC="${MPIINC:-/opt/conda/include}"
mkdir -p "$OUT"
FFLAGS="-O2 -fPIC -I$MPIINC -J$OUT -I$OUT -w"
"$FC" $FFLAGS -c "$REF/declaration.f90" -o "$OUT/declaration.o"
for f in littleharsh tridLU_3D_IB y_grid_canopy geom_sym_canopy spectra stats start; do
  "$FC" $FFLAGS -c "$REF/$f.f90" -o "$OUT/$f.o"
done
"$FC" $FFLAGS -c "$HERE/refdrv.f90" -o "$OUT/refdrv.o"
"$FC" -shared -fPIC -o "$OUT/libref.so" "$OUT/declaration.o" "$OUT/littleharsh.o" \
  "$OUT/tridLU_3D_IB.o" "$OUT/y_grid_canopy.o" "$OUT/geom_sym_canopy.o" "$OUT/spectra.o" "$OUT/stats.o" \
  "$OUT/start.o" "$OUT/refdrv.o"
"$FC" -O2 -fPIC -shared -w -o "$OUT/libfftpack.so" "$REF/rft_buff.f" "$REF/cft_buff.f"
echo "built $OUT/libref.so $OUT/libfftpack.so"
