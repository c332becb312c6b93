#!/usr/bin/env bash
# Build the reference's sequential Fortran path (the compiled oracle) for one
# vertical resolution, from the sources where they lie under /root/reference.
#
#   oracle/ref/build_ref.sh NN [Lx Ly H TopBC BotBC XPB YPB XBC YBC]
#
# Output goes only to oracle/_ref/nn<NN>/ (git-ignored):
#   inc/    configured copies of src/Include/{cppdefs.h,common.mpf}: the two
#           files the reference expects a user to edit per case (layer count
#           NN, box size XDFin/YDFin/sigFin, TopBC/BotBC, the horizontal
#           periodicity XPB/YPB and wall kinds XBC/YBC, KeyParallel off for
#           the sequential build, KeyLES on);
#   obj/    objects and .mod files of every non-MPI reference source;
#   nsx_ref_driver   the reference objects linked with the driver program in
#           oracle/ref/nsx_ref_driver.F90 (written for this repository).
# Nothing under /root/reference is modified; no reference source is copied
# into the repository.
set -euo pipefail
NN=${1:?usage: build_ref.sh NN [Lx Ly H TopBC BotBC XPB YPB XBC YBC]}
LX=${2:-6.283185307179586}
LY=${3:-3.141592653589793}
HH=${4:-2.0}
TOPBC=${5:-0}
BOTBC=${6:-0}
XPB=${7:-1}
YPB=${8:-1}
XBC=${9:-0}
YBC=${10:-1}
REF=${NSX_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
# a box other than the default gets a directory of its own, and so do
# boundary conditions other than the channel's
if [ "$LX" = "6.283185307179586" ] && [ "$LY" = "3.141592653589793" ]; then
  OUT=$HERE/../_ref/nn$NN
else
  OUT=$HERE/../_ref/nn${NN}_box${LX}x${LY}
fi
if [ "$TOPBC$BOTBC$XPB$YPB$XBC$YBC" != "001101" ]; then
  OUT=${OUT}_bc${TOPBC}${BOTBC}${XPB}${YPB}${XBC}${YBC}
fi
FC=${FC:-/opt/rocm/bin/amdflang}
FFLAGS=${FFLAGS:--O2}

[ -d "$REF/src" ] || { echo "reference sources not found at $REF" >&2; exit 3; }
mkdir -p "$OUT/inc" "$OUT/obj"
tr -d '\r' < "$REF/src/Include/cppdefs.h" \
  | sed -e 's/^#\( *\)define KeyParallel/!#\1define KeyParallel/' \
        -e 's/^!#\( *\)define KeyLES$/#\1define KeyLES/' > "$OUT/inc/cppdefs.h"
tr -d '\r' < "$REF/src/Include/common.mpf" \
  | sed -e "s/^\( *real\*8, parameter :: sigFin =\).*/\1 ${HH}d0/" \
        -e "s/^\( *real\*8, parameter :: XDFin =\).*/\1 ${LX}d0/" \
        -e "s/^\( *real\*8, parameter :: YDFin =\).*/\1 ${LY}d0/" \
        -e "s/^\( *integer,parameter :: TopBC =\).*/\1 ${TOPBC}/" \
        -e "s/^\( *integer,parameter :: BotBC =\).*/\1 ${BOTBC}/" \
        -e "s/^\( *integer,parameter :: XPB =\).*/\1 ${XPB}/" \
        -e "s/^\( *integer,parameter :: YPB =\).*/\1 ${YPB}/" \
        -e "s/^\( *integer,parameter :: XBC =\).*/\1 ${XBC}/" \
        -e "s/^\( *integer,parameter :: YBC =\).*/\1 ${YBC}/" \
        -e "s/^\( *integer,parameter:: NN =\).*/\1 ${NN}/" > "$OUT/inc/common.mpf"
grep -q "define KeyParallel" "$OUT/inc/cppdefs.h" && ! grep -q "^#.*define KeyParallel" "$OUT/inc/cppdefs.h"
grep -q "^#.*define KeyLES$" "$OUT/inc/cppdefs.h"
grep -q "NN = ${NN}$" "$OUT/inc/common.mpf"
grep -q "XPB = ${XPB}$" "$OUT/inc/common.mpf" && grep -q "YPB = ${YPB}$" "$OUT/inc/common.mpf"
grep -q "XBC = ${XBC}$" "$OUT/inc/common.mpf" && grep -q "YBC = ${YBC}$" "$OUT/inc/common.mpf"

# module order of the reference Makefile, MPI sources and the PROGRAM left out
SRCS="Main/module_variables Main/module_geometry Main/module_interfaces LES/module_les Stats/module_stats
Main/hydro Main/input Main/input_parameters Main/calcul_geometry Main/initial Main/user_initial Main/cfl
Stats/stats Main/bicgstab Save/SavetecVertex Stats/SavetecStats Save/SaveParaviewVertex Save/SavetecCenter
Save/SavetecVertexCenter LoopUpdate/LoopTimeInitial LoopUpdate/LoopTimeUpdate LoopUpdate/LoopRKInitial
LoopUpdate/LoopRKUpdate LoopUpdate/LoopRKFinal Stats/LoopStats Interpolation/interpolation2D
Interpolation/interpolation3D Interpolation/interpolationFace Interpolation/interpolationU Gradient/gradientLSM
Gradient/gradientLSM2D Gradient/gradientFVM Gradient/gradientP AdvectionDiffusion/advection3D
AdvectionDiffusion/diffusion3D AdvectionDiffusion/PDE AdvectionDiffusion/source3D AdvectionDiffusion/newu3D
AdvectionDiffusion/AdvDiffNS AdvectionDiffusion/AdvDiffSolve BoundaryConditions/BC BoundaryConditions/BCoflow
BoundaryConditions/update_bc LES/sgm_les Main/Functions"
OBJS=""
for s in $SRCS; do
  o="$OUT/obj/$(basename "$s").o"
  if [ ! -f "$o" ] || [ "$REF/src/$s.F90" -nt "$o" ] || [ "$OUT/inc/common.mpf" -nt "$o" ]; then
    (cd "$OUT/obj" && "$FC" $FFLAGS -I "$OUT/inc" -J "$OUT/obj" -c "$REF/src/$s.F90" -o "$o")
  fi
  OBJS="$OBJS $o"
done
(cd "$OUT/obj" && "$FC" $FFLAGS -I "$OUT/inc" -J "$OUT/obj" -c "$HERE/nsx_ref_driver.F90" -o "$OUT/obj/nsx_ref_driver.o")
"$FC" $FFLAGS -o "$OUT/nsx_ref_driver" "$OUT/obj/nsx_ref_driver.o" $OBJS -lm
echo "built $OUT/nsx_ref_driver"
