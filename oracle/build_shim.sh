#!/bin/sh
# Builds oracle/_ref/shim/shim_driver: this directory's golden-vector driver
# (ref_driver.f90) and the reference's own MC.f90, runmc.f90 and the other
# modules they use, compiled unchanged from where they lie (REF, default
# /root/reference; read, never copied), with the reference's FourierGrid,
# SumSinCosKR, RhoSquared, EwaldSumKSpace, ForceKSpace, RealSumLocal,
# EwaldSumRealSpace and EwaldSumTails modules replaced by the drop-in modules
# of fortran/shim/, linked to yamc_amd/libyamc_ewald.so.  Its output is
# compared with the golden fixtures the pure reference build wrote: the
# reference's Monte Carlo code running on the device library.
# Output goes to oracle/_ref/shim only (not committed).  Test infrastructure.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(dirname "$HERE")
REF=${REF:-/root/reference}
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
OUT=${OUT:-"$HERE/_ref/shim"}
LIB="$ROOT/yamc_amd/libyamc_ewald.so"
[ -d "$REF" ] || { echo "build_shim.sh: $REF not present, skipping"; exit 0; }
[ -x "$FC" ] || { echo "build_shim.sh: $FC not found, skipping"; exit 0; }
[ -f "$LIB" ] || { echo "build_shim.sh: $LIB not built (python -m yamc_amd.build)"; exit 1; }
mkdir -p "$OUT"
cd "$OUT"
SHIM="FourierGrid SumSinCosKR RhoSquared EwaldSumKSpace ForceKSpace RealSumLocal EwaldSumRealSpace EwaldSumTails"
# dependency order; the shim modules take their reference counterparts' places
MODS="SystemSettings constants error string io matrix3x3 geometry random Functions parameters \
Molecule MoleculeHandler composition AtomicData LJTypes yamc_ewald yamc_shim_core $SHIM \
EwaldSumExternal MonteCarloMove MoleculeTable BiasedRandom scale_box MC FloatingPoint runmc"
OBJS=""
for m in $MODS; do
  case " yamc_ewald " in *" $m "*) src="$ROOT/fortran/$m.f90";; *)
    case " yamc_shim_core $SHIM " in *" $m "*) src="$ROOT/fortran/shim/$m.f90";; *) src="$REF/$m.f90";; esac;;
  esac
  "$FC" -O2 -c "$src" -o "$m.o"
  OBJS="$OBJS $m.o"
done
"$FC" -O2 -c "$HERE/ref_driver.f90" -o ref_driver.o
# the library is found next to the repository's tree wherever it is copied
"$FC" -O2 -o shim_driver ref_driver.o $OBJS -L"$ROOT/yamc_amd" -lyamc_ewald -Wl,-rpath,'$ORIGIN/../../../yamc_amd' -Wl,-rpath,"$ROOT/yamc_amd"
echo "$OUT/shim_driver"
