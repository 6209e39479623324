#!/bin/bash
# Compiles the reference's own random-number and sort routines (accesories.f90,
# with global.f90 for MODULE PARAMS) from /root/reference into oracle/_ref/, so the
# oracle's restatement of ranff / ranpp / ranff2 can be checked bit for bit.
# The reference's build flag -r8 (Makefile:6) becomes flang's -fdefault-real-8.
# Only these two files compile on their own: flowlib.f90 needs Intel's IFPORT,
# partlib.f90 uses Intel-only variable formats and mpifft.f90 needs FFTW 2.1.5.
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/accesories.f90" ]; then
  echo "reference sources not present; skipping oracle/_ref" >&2
  exit 0
fi
FC=${FC:-/opt/rocm/bin/amdflang}
mkdir -p "$OUT/mod"
"$FC" -O0 -fPIC -fdefault-real-8 -J "$OUT/mod" -c "$REF/global.f90" -o "$OUT/global.o"
"$FC" -O0 -fPIC -fdefault-real-8 -J "$OUT/mod" -c "$REF/accesories.f90" -o "$OUT/accesories.o"
"$FC" -shared -fPIC "$OUT/global.o" "$OUT/accesories.o" -o "$OUT/libref_acc.so"
rm -f "$OUT/global.o" "$OUT/accesories.o"
echo "built $OUT/libref_acc.so"
