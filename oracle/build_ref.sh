#!/bin/bash
# Builds oracle/_ref/ref_harness from the reference's own Fortran sources,
# compiled where they lie under /root/reference with flang and the image's
# MPICH headers/libraries.  Test infrastructure only.
#
# Built: every module the substep needs except the Poisson transform modules
# (fastfouriertransform.F90, poisson_solve.F90), which need FFTW (absent from
# this image).  part_tools.F90 is rejected by flang at line 454 (a formatted
# write), so only its leading type module (part_type, lines 1-50, needed by
# `type run`) is compiled, taken verbatim by line range into a temporary
# directory outside the repository.
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
[ -d "$REF" ] || { echo "reference not present: $REF" >&2; exit 3; }
MPIINC=${MPIINC:-/opt/conda/include}
MPILIB=${MPILIB:-/opt/conda/lib}
FC=${FC:-flang}
TMP=$(mktemp -d /tmp/vring_ref.XXXXXX)
trap 'rm -rf "$TMP"' EXIT
mkdir -p "$OUT"
cd "$TMP"
FFLAGS="-cpp -O2 -I$MPIINC"
sed -n '/^module part_type/,/^end module part_type/p' "$REF/part_tools.F90" > part_type.F90
SRCS="gen_tools_swirl cons_tools comm_tools para_tools mesh_tools solp_tools solu_toolsII"
for f in $SRCS; do $FC $FFLAGS -c "$REF/$f.F90" -o $f.o 2>/dev/null; done
$FC $FFLAGS -c part_type.F90 -o part_type.o 2>/dev/null
for f in run tripar momentum_coef momentum_terms_swirl step communicate divergence bessel; do
  $FC $FFLAGS -c "$REF/$f.F90" -o $f.o 2>/dev/null
done
$FC $FFLAGS -c "$HERE/ref_harness.F90" -o ref_harness.o 2>/dev/null
$FC -o "$OUT/ref_harness" ref_harness.o gen_tools_swirl.o cons_tools.o comm_tools.o para_tools.o \
   mesh_tools.o solp_tools.o solu_toolsII.o part_type.o run.o tripar.o momentum_coef.o \
   momentum_terms_swirl.o step.o communicate.o divergence.o bessel.o \
   -L"$MPILIB" -Wl,-rpath,"$MPILIB" -lmpifort -lmpi
echo "built $OUT/ref_harness"
