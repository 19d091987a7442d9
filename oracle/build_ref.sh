#!/usr/bin/env bash
# ORACLE -- test infrastructure only.  Compiles the reference's own Fortran leaf
# modules, where they lie under /root/reference, with AMD flang, plus our
# bind(C) shim, into oracle/_ref/libiceref.so.  Nothing is copied: the sources
# are read in place; objects and module files go to oracle/_ref/.
# The modules that need ORION or FiNeR (block-level ones) are not built.
set -euo pipefail
REF=${ICE_REFERENCE:-/root/reference}/src/lib
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
FC=${FC_REF:-/opt/rocm/lib/llvm/bin/flang}
[ -d "$REF" ] || { echo "reference sources not present; oracle/_ref not built"; exit 0; }
mkdir -p "$OUT/obj"
FLAGS="-O2 -fPIC -cpp -ffp-contract=off -module-dir $OUT/obj"
SRCS="base/Parameters_m.f90 base/Finite_m.f90 base/Global_m.f90 physics/Lib_Evaporation.f90
config/Config_Types_m.f90 physics/Lib_Properties.f90 physics/Lib_Solidification.f90
physics/Lib_Model_MK.f90 physics/Lib_Model_IG.f90 physics/Lib_Model_AG.f90 physics/Lib_Model_Solid.f90
physics/Lib_Drag.f90 physics/Lib_Heat.f90 physics/Lib_Model.f90 numerics/space/Lib_Limiters.f90
numerics/time/explicit/Lib_RK.f90 numerics/space/Lib_Shock_Detector.f90 numerics/space/Lib_Reconstruction.f90
numerics/fluxes/riemann/Lib_Riemann.f90 physics/Lib_Drop_Surface.f90"
OBJS=""
for s in $SRCS; do
  o=$OUT/obj/$(basename "${s%.f90}").o
  if [ ! -f "$o" ] || [ "$REF/$s" -nt "$o" ]; then
    $FC $FLAGS -c "$REF/$s" -o "$o"
  fi
  OBJS="$OBJS $o"
done
$FC $FLAGS -c "$HERE/ref_shim.f90" -o "$OUT/obj/ref_shim.o"
$FC -shared -o "$OUT/libiceref.so.tmp" $OBJS "$OUT/obj/ref_shim.o"
mv "$OUT/libiceref.so.tmp" "$OUT/libiceref.so"
echo "built $OUT/libiceref.so"
