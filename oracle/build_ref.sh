#!/bin/bash
# Compile the reference's linear-solve sources (unmodified, read in place) and
# link them with ref_solve_driver.f90.  Outputs only under oracle/_ref/.
set -euo pipefail
cd "$(dirname "$0")"
: "${REF:=/root/reference}"
: "${FC:=amdflang}"
: "${FFLAGS:=-O2 -ffixed-line-length-none -fno-automatic -fopenmp -module-dir _ref/mod -I _ref/mod}"
if [ ! -d "$REF/SolveEquations" ]; then
  echo "reference sources not present at $REF; skipping reference build" >&2
  exit 0
fi
mkdir -p _ref/obj _ref/mod
SRCS=()
for f in "$REF"/DeclareVariables/*; do SRCS+=("$f"); done
SRCS+=("$REF/ProblemDesciption/User0.F90")
for f in "$REF"/SolveEquations/AlgebraicSolvers/* "$REF"/SolveEquations/AlgebraicMultiGrid/* \
         "$REF"/SolveEquations/FlowInArteries/*; do SRCS+=("$f"); done
for f in "$REF"/Utilities/*; do
  case "$(basename "$f")" in TimeStampRoutine.for) ;; *) SRCS+=("$f") ;; esac
done
OBJS=()
for f in "${SRCS[@]}"; do
  b=$(basename "$f"); o="_ref/obj/${b%.*}.o"
  if [ ! -f "$o" ] || [ "$f" -nt "$o" ]; then $FC $FFLAGS -c "$f" -o "$o"; fi
  OBJS+=("$o")
done
$FC $FFLAGS -c ref_solve_driver.f90 -o _ref/ref_solve_driver.o
$FC -fopenmp -o _ref/ref_solve.x _ref/ref_solve_driver.o "${OBJS[@]}"
echo "built _ref/ref_solve.x"

# The gradient limiter and the high-resolution convection correction with
# their limiter functions, linked with ref_conv_driver.f90.
CONV_OBJS=()
for f in "$REF"/SolveEquations/GradientFiles/LimitGradients.for \
         "$REF"/SolveEquations/Convection/AssembleHRSchemeCorrection.for \
         "$REF"/SolveEquations/Convection/LimitersTVDType.for \
         "$REF"/SolveEquations/Convection/LimitersNVFType.for; do
  b=$(basename "$f"); o="_ref/obj/${b%.*}.o"
  if [ ! -f "$o" ] || [ "$f" -nt "$o" ]; then $FC $FFLAGS -c "$f" -o "$o"; fi
  CONV_OBJS+=("$o")
done
$FC $FFLAGS -c ref_conv_driver.f90 -o _ref/ref_conv_driver.o
$FC -fopenmp -o _ref/ref_conv.x _ref/ref_conv_driver.o "${CONV_OBJS[@]}" "${OBJS[@]}"
echo "built _ref/ref_conv.x"

# The reference's physics sources (geometry processing, assembly, gradients,
# interpolation, continuity, boundary updates, turbulence support) in an
# archive, so that the linker takes only what ref_simple_driver.f90 reaches.
# The mesh reader and the storage allocator are left out (they need source
# patches); the driver installs their results itself.
mkdir -p _ref/simple_obj
AR_SRCS=()
AR_OBJS=()
for d in Assemble Buoyancy ContinuityEquation Convection Diffusion Energy FalseTransient GradientFiles \
         InitializeFluxesAndCoefficients Interpolation Momentum Periodicity PhysicalProperties \
         SetBoundaryPatches SurfaceTension Transient UnderRelaxVariables UpdateVariablesAtBoundaries \
         UpdateVariablesFromBoundaryConditions NormalAndTangentialVelocities Tensors FarFieldVariables \
         WriteAndReadSolution; do
  for f in "$REF/SolveEquations/$d"/*; do AR_SRCS+=("$f"); done
done
AR_SRCS+=("$REF/SolveEquations/GridAndGeometry/ProcessGeometryRoutine.for")
while IFS= read -r f; do AR_SRCS+=("$f"); done < <(find "$REF/Turbulence" -type f \( -name '*.for' -o -name '*.f90' -o -name '*.F90' \) | sort)
for f in "${AR_SRCS[@]}"; do
  b=$(basename "$f"); o="_ref/simple_obj/${b%.*}.o"
  if [ -f "_ref/obj/${b%.*}.o" ]; then continue; fi
  if [ ! -f "$o" ] || [ "$f" -nt "$o" ]; then
    $FC $FFLAGS -c "$f" -o "$o" 2> "_ref/simple_obj/${b%.*}.err" || { rm -f "$o"; continue; }
  fi
  AR_OBJS+=("$o")
done
rm -f _ref/libref_simple.a
ar rcs _ref/libref_simple.a "${AR_OBJS[@]}" "${CONV_OBJS[@]}"
$FC $FFLAGS -c ref_simple_driver.f90 -o _ref/ref_simple_driver.o
$FC -fopenmp -o _ref/ref_simple.x _ref/ref_simple_driver.o "${OBJS[@]}" _ref/libref_simple.a
echo "built _ref/ref_simple.x"

# The same driver linked with the device drop-in SolveEquation (the Fortran
# shim of the product) instead of SolveAlegbraicEquations.for.
LIBDIR="$(cd .. && pwd)/flowsolversourcecode_amd"
if [ -f "$LIBDIR/libfsamd.so" ]; then
  $FC $FFLAGS -c "$LIBDIR/fortran/fsamd_iface.f90" -o _ref/fsamd_iface.o
  $FC $FFLAGS -c "$LIBDIR/fortran/SolveEquationDevice.f90" -o _ref/SolveEquationDevice.o
  OBJS2=()
  for o in "${OBJS[@]}"; do
    case "$(basename "$o")" in SolveAlegbraicEquations.o) ;; *) OBJS2+=("$o") ;; esac
  done
  $FC -fopenmp -o _ref/ref_solve_device.x _ref/ref_solve_driver.o _ref/fsamd_iface.o \
      _ref/SolveEquationDevice.o "${OBJS2[@]}" -L"$LIBDIR" -lfsamd -Wl,-rpath,"\$ORIGIN/../../flowsolversourcecode_amd"
  echo "built _ref/ref_solve_device.x"
fi
