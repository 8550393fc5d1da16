#!/bin/sh
# Compiles the reference YAMC Fortran modules of the hot path where they lie
# (REF, default /root/reference; sources are read, never copied) plus this
# directory's golden-vector driver into oracle/_ref/ref_driver.
# Output goes to oracle/_ref only (not committed).  Test infrastructure.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
REF=${REF:-/root/reference}
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
OUT="$HERE/_ref"
[ -d "$REF" ] || { echo "build_ref.sh: $REF not present, skipping"; exit 0; }
[ -x "$FC" ] || { echo "build_ref.sh: $FC not found, skipping"; exit 0; }
mkdir -p "$OUT"
cd "$OUT"
# dependency order of the modules that the driver reaches
MODS="SystemSettings constants error string io matrix3x3 geometry random Functions parameters \
Molecule MoleculeHandler composition AtomicData LJTypes FourierGrid SumSinCosKR RhoSquared \
EwaldSumKSpace ForceKSpace RealSumLocal EwaldSumRealSpace EwaldSumTails EwaldSumExternal \
MonteCarloMove MoleculeTable BiasedRandom scale_box MC FloatingPoint runmc"
OBJS=""
for m in $MODS; do
  if [ ! -f "$m.o" ] || [ "$REF/$m.f90" -nt "$m.o" ]; then
    "$FC" -O2 -c "$REF/$m.f90" -o "$m.o"
  fi
  OBJS="$OBJS $m.o"
done
"$FC" -O2 -c "$HERE/ref_driver.f90" -o ref_driver.o
"$FC" -O2 -o ref_driver ref_driver.o $OBJS
echo "$OUT/ref_driver"
