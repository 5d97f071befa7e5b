#!/bin/sh
# Builds the reference's explicit step (Fortran sources read in place under
# $MPM3D_REF/src, default /root/reference/src) together with the dump driver
# oracle/ref_driver.f90 into oracle/_ref/ref_driver.
#
# Flags follow the parity-oracle choice: -O2 -ffp-contract=off (no FMA
# contraction, no fast-math), single thread.
#
# One source line cannot be compiled by flang as written: Grid.f90:210 uses an
# INTEGER array element as an IF condition.  The recipe compiles a copy of
# Grid.f90 with that single line rewritten to the explicit test "/=0" (the
# form the same array is tested with at update_step.f90:1447).  The copy lives
# in a temporary directory outside the repository and is deleted afterwards;
# only objects, module files and the executable are written, into oracle/_ref/.
# The line only matters for SGMP particles next to fixed/symmetric faces.
set -eu
REF=${MPM3D_REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
FC=${FC:-flang}
FFLAGS=${FFLAGS:--O2 -ffp-contract=off}

if [ ! -f "$REF/src/update_step.f90" ]; then
  echo "build_ref.sh: reference sources not found under $REF/src" >&2
  exit 3
fi

TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT INT TERM
mkdir -p "$OUT"
sed '210s/if(p_type(i))then/if(p_type(i)\/=0)then/' "$REF/src/Grid.f90" > "$TMP/Grid.f90"
if ! grep -q 'if(p_type(i)/=0)then' "$TMP/Grid.f90"; then
  echo "build_ref.sh: expected line Grid.f90:210 not found" >&2
  exit 4
fi

cd "$OUT"
for f in "$REF/src/FFI.f90" "$REF/src/Material.f90" "$REF/src/Particle.f90" \
         "$TMP/Grid.f90" "$REF/src/DataOut.f90" "$REF/src/DataIn.f90" \
         "$REF/src/Constitution.f90" "$REF/src/update_step.f90"; do
  b=$(basename "$f" .f90)
  $FC $FFLAGS -w -c "$f" -o "$OUT/$b.o"
done
$FC $FFLAGS -w -c "$REF/src/MPM3D.f90" -o "$OUT/MPM3D.o"
# the reference program itself (its own main loop, writers and step)
$FC $FFLAGS -o "$OUT/mpm3d_ref" MPM3D.o FFI.o Material.o Particle.o Grid.o DataOut.o DataIn.o Constitution.o \
    update_step.o
echo "built $OUT/mpm3d_ref"
$FC $FFLAGS -c "$HERE/ref_driver.f90" -o "$OUT/ref_driver.o"
$FC $FFLAGS -o "$OUT/ref_driver" "$OUT/ref_driver.o" FFI.o Material.o Particle.o \
    Grid.o DataOut.o DataIn.o Constitution.o update_step.o
echo "built $OUT/ref_driver"

# Drop-in check: the reference's own main program (MPM3D.f90) and input
# phase linked with mpm3d-f90_amd/fortran/update_step_hip.f90 in place of
# update_step.o, driving libmpm3d_hip.so.  Built only when the library exists.
LIBDIR="$HERE/../mpm3d-f90_amd"
if [ -f "$LIBDIR/libmpm3d_hip.so" ]; then
  $FC $FFLAGS -c "$LIBDIR/fortran/mpm_hip.f90" -o "$OUT/mpm_hip.o"
  $FC $FFLAGS -w -c "$LIBDIR/fortran/update_step_hip.f90" -o "$OUT/update_step_hip.o"
  $FC $FFLAGS -o "$OUT/mpm3d_dropin" MPM3D.o update_step_hip.o mpm_hip.o FFI.o Material.o Particle.o \
      Grid.o DataOut.o DataIn.o Constitution.o -L"$LIBDIR" -lmpm3d_hip -Wl,-rpath,"$LIBDIR"
  echo "built $OUT/mpm3d_dropin"
fi
