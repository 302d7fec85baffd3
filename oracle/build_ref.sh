#!/bin/bash
# Builds oracle/_ref/libtacs_ref.so from the reference's element-level sources
# where they lie under /root/reference, plus oracle/ref_driver.cpp.
# TACSAssembler.cpp is not built: it includes metis.h, which this image lacks.
# Uses the MPICH headers/library and MKL LAPACK installed under /opt/conda.
set -e
REF=${TACS_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
mkdir -p "$OUT"
INC="-I$REF/src -I$REF/src/bpmat -I$REF/src/elements -I$REF/src/elements/basis \
 -I$REF/src/elements/shell -I$REF/src/elements/a2d -I$REF/src/constitutive -I$REF/src/functions \
 -I/opt/conda/include"
SRCS="src/TACSObject.cpp src/TacsUtilities.cpp
 src/elements/TACSElement.cpp src/elements/TACSElementTypes.cpp src/elements/TACSElementVerification.cpp
 src/elements/TACSElement3D.cpp src/elements/TACSLinearElasticity.cpp src/elements/TACSTraction3D.cpp
 src/elements/TACSPressure3D.cpp src/elements/TACSInertialForce3D.cpp src/elements/TACSCentrifugalForce3D.cpp
 src/elements/basis/TACSElementBasis.cpp src/elements/basis/TACSHexaBasis.cpp
 src/elements/basis/TACSTensorProductBasisImpl.cpp
 src/constitutive/TACSConstitutive.cpp src/constitutive/TACSMaterialProperties.cpp
 src/constitutive/TACSSolidConstitutive.cpp src/constitutive/TACSShellConstitutive.cpp
 src/constitutive/TACSIsoShellConstitutive.cpp src/constitutive/TACSBeamConstitutive.cpp
 src/bpmat/BCSRMatMult.cpp src/bpmat/BCSRMatMult3.cpp src/bpmat/BCSRMatMult6.cpp"
OBJS=""
for s in $SRCS; do
  o=$OUT/$(basename "${s%.cpp}").o
  if [ ! -f "$o" ] || [ "$REF/$s" -nt "$o" ]; then
    g++ -O2 -fPIC -std=c++11 -w $INC -c "$REF/$s" -o "$o"
  fi
  OBJS="$OBJS $o"
done
g++ -O2 -fPIC -std=c++11 -w $INC -c "$HERE/ref_driver.cpp" -o "$OUT/ref_driver.o"
g++ -shared -o "$OUT/libtacs_ref.so" $OBJS "$OUT/ref_driver.o" \
  -L/opt/conda/lib -Wl,-rpath,/opt/conda/lib -lmpi -lmkl_rt
echo "built $OUT/libtacs_ref.so"

# The same sources with TACS_USE_COMPLEX (complex-step element derivatives)
# plus oracle/ref_driver_cplx.cpp -> oracle/_ref/cplx/libtacs_ref_cplx.so
OUTC=$OUT/cplx
mkdir -p "$OUTC"
OBJSC=""
for s in $SRCS; do
  o=$OUTC/$(basename "${s%.cpp}").o
  if [ ! -f "$o" ] || [ "$REF/$s" -nt "$o" ]; then
    g++ -O2 -fPIC -std=c++11 -w -DTACS_USE_COMPLEX $INC -c "$REF/$s" -o "$o"
  fi
  OBJSC="$OBJSC $o"
done
g++ -O2 -fPIC -std=c++11 -w -DTACS_USE_COMPLEX $INC -c "$HERE/ref_driver_cplx.cpp" \
  -o "$OUTC/ref_driver_cplx.o"
g++ -shared -o "$OUTC/libtacs_ref_cplx.so" $OBJSC "$OUTC/ref_driver_cplx.o" \
  -L/opt/conda/lib -Wl,-rpath,/opt/conda/lib -lmpi -lmkl_rt
echo "built $OUTC/libtacs_ref_cplx.so"
