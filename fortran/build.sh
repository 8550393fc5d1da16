#!/bin/sh
# Builds the Fortran binding module (fortran/_build/yamc_ewald.mod, .o) and
# the Fortran host fortran/mc_host against yamc_amd/libyamc_ewald.so.
# Skips (exit 0) when no Fortran compiler is installed.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
[ -x "$FC" ] || { echo "fortran/build.sh: $FC not found, skipping"; exit 0; }
mkdir -p "$HERE/_build"
cd "$HERE"
"$FC" -O2 -c yamc_ewald.f90 -module-dir _build -o _build/yamc_ewald.o
"$FC" -O2 -I _build -c mc_host.f90 -module-dir _build -o _build/mc_host.o
"$FC" -o mc_host _build/mc_host.o _build/yamc_ewald.o -L../yamc_amd -lyamc_ewald -Wl,-rpath,'$ORIGIN/../yamc_amd'
echo "$HERE/mc_host"
