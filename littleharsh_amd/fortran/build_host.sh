#!/bin/bash
# Build the reference Fortran host's set-up for one rank (lh_host_setup): the reference's own
# set-up sources compiled where they lie, with the single-rank proc_lims_planes and getbounds
# of lh_single_rank.f90 linked in place of the two routines of start.f90 (their symbols in
# the start.f90 object are made weak), and lh_host.f90's hand-over to lh_create.  The objects
# form a shared library bound lazily: routines that the set-up never reaches, among them the
# ones that call the absent FOU3D.f90, keep those calls unresolved and are never entered.
# Needs the reference sources (LH_REFERENCE, default /root/reference) and MPICH.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${LH_REFERENCE:-/root/reference}"
FC="${FC:-/opt/rocm/bin/amdflang}"
MPIINC="${MPIINC:-/opt/conda/include}"
MPILIB="${MPILIB:-/opt/conda/lib}"
OUT="$HERE/_build/host"
mkdir -p "$OUT"
FFLAGS="-O2 -fPIC -I$MPIINC -J$OUT -I$OUT -w"
"$FC" $FFLAGS -c "$REF/declaration.f90" -o "$OUT/declaration.o"
for f in start y_grid_canopy geom_sym_canopy tridLU_3D_IB; do
  "$FC" $FFLAGS -c "$REF/$f.f90" -o "$OUT/$f.o"
done
"$FC" -O2 -fPIC -w -c "$REF/rft_buff.f" -o "$OUT/rft_buff.o"
"$FC" -O2 -fPIC -w -c "$REF/cft_buff.f" -o "$OUT/cft_buff.o"
objcopy --weaken-symbol=proc_lims_planes_ --weaken-symbol=getbounds_ "$OUT/start.o" "$OUT/start_1rank.o"
"$FC" $FFLAGS -c "$HERE/lh_amd.f90" -o "$OUT/lh_amd.o"
"$FC" $FFLAGS -c "$HERE/lh_single_rank.f90" -o "$OUT/lh_single_rank.o"
"$FC" $FFLAGS -c "$HERE/lh_host.f90" -o "$OUT/lh_host.o"
"$FC" $FFLAGS -c "$HERE/lh_host_setup.f90" -o "$OUT/lh_host_setup.o"
"$FC" -shared -fPIC -o "$OUT/libhostsetup.so" "$OUT/lh_host.o" "$OUT/lh_amd.o" "$OUT/lh_single_rank.o" \
  "$OUT/start_1rank.o" "$OUT/y_grid_canopy.o" "$OUT/geom_sym_canopy.o" "$OUT/tridLU_3D_IB.o" \
  "$OUT/declaration.o" "$OUT/rft_buff.o" "$OUT/cft_buff.o" -L"$MPILIB" -lmpifort -lmpi -Wl,-rpath,"$MPILIB"
"$FC" -o "$HERE/lh_host_setup" "$OUT/lh_host_setup.o" -L"$OUT" -lhostsetup -Wl,--allow-shlib-undefined \
  -Wl,-z,lazy -Wl,-rpath,'$ORIGIN/_build/host' -L"$MPILIB" -lmpifort -lmpi -Wl,-rpath,"$MPILIB"
echo "built $HERE/lh_host_setup"
