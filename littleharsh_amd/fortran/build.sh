#!/bin/bash
# Build the Fortran interface module and the Fortran host demo against liblh_amd.so.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
FC="${FC:-/opt/rocm/bin/amdflang}"
OUT="$HERE/_build"
mkdir -p "$OUT"
"$FC" -O2 -J"$OUT" -c "$HERE/lh_amd.f90" -o "$OUT/lh_amd.o"
"$FC" -O2 -I"$OUT" -c "$HERE/lh_demo.f90" -o "$OUT/lh_demo.o"
"$FC" -o "$HERE/lh_demo" "$OUT/lh_demo.o" "$OUT/lh_amd.o" -L"$HERE/.." -llh_amd \
  -Wl,-rpath,'$ORIGIN/..'
echo "built $HERE/lh_demo"
