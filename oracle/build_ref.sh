#!/usr/bin/env bash
# build_ref.sh -- compile the reference's Fortran LBL path from its sources, in place.
#
# TEST INFRASTRUCTURE ONLY.  The reference (Razumovskyy/MARFA) is plain free-form Fortran
# with no generated code and no external libraries, so it is compiled here directly
# with amdflang in module dependency order (its own fpm build is not used; fpm is not in
# the image).  Sources are read where they lie; every output (.o, .mod, binaries) goes to
# oracle/_ref/, which is git-ignored.  -ffp-contract=off keeps a*b+c unfused so that the
# C restatement (built with the same setting) can replay the reference bit for bit.
#
# Outputs:
#   oracle/_ref/marfa_ref   the reference's own program (app/main.f90)
#   oracle/_ref/ref_probe   oracle/ref_probe.f90 over the reference modules
#   oracle/_ref/ref_driver  oracle/ref_driver.f90 over the reference modules
#   oracle/_ref/processParFile  the reference's .par preprocessing program
#                               (app/processParFile.f90)
set -euo pipefail
REF=${MARFA_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)
OUT="$HERE/_ref"
FC=${FC:-amdflang}
FFLAGS="-O2 -ffp-contract=off"

if [ ! -d "$REF/src" ]; then
    echo "build_ref.sh: reference sources not present at $REF; keeping prebuilt $OUT" >&2
    exit 0
fi
mkdir -p "$OUT"
cd "$OUT"
MODS="Constants Interfaces Atmsosphere MolarMasses ChiFactors Spectroscopy Shapes Grids LineGridCalc LBL EnvironmentSetup"
OBJS=""
for m in $MODS; do
    $FC $FFLAGS -c "$REF/src/$m.f90" -o "$m.o" -module-dir "$OUT"
    OBJS="$OBJS $m.o"
done
$FC $FFLAGS -c "$REF/app/main.f90" -o main.o -module-dir "$OUT"
$FC $FFLAGS -o marfa_ref $OBJS main.o
$FC $FFLAGS -c "$HERE/ref_probe.f90" -o ref_probe.o -module-dir "$OUT" -I "$OUT"
$FC $FFLAGS -o ref_probe $OBJS ref_probe.o
$FC $FFLAGS -c "$HERE/ref_driver.f90" -o ref_driver.o -module-dir "$OUT" -I "$OUT"
$FC $FFLAGS -o ref_driver $OBJS ref_driver.o
$FC $FFLAGS -o processParFile "$REF/app/processParFile.f90" -module-dir "$OUT"
echo "build_ref.sh: built $OUT/marfa_ref $OUT/ref_probe $OUT/ref_driver $OUT/processParFile"
