#!/bin/bash
# Compiles the reference LAP CPU path (serial simple_blas build and OpenMP
# build) from its own source files, read in place, into oracle/_ref/.  Used
# only to validate the oracle and to produce golden fixtures; nothing is
# copied out of the reference tree.  Optional second step: when the HIP
# library lap_amd/liblapgfx.so exists, link the reference's own host solver
# and driver (cg_driver2.c, cg.c, prec.c, GS.c, ...) built with -DHIP=1
# against it, which is the drop-in link check.
set -euo pipefail
REF=${LAP_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
if [ ! -f "$REF/cg.c" ]; then
  echo "build_ref: reference sources not present at $REF; nothing to build"
  exit 0
fi
mkdir -p "$OUT"
CC=${CC:-gcc}
SOLVER="blas.c GS.c it_jacobi.c line_jacobi.c prec.c cg.c io_utils.c"
CPU_DEFS="-DV100=0 -DNOACC=1 -DCUDA=0 -DOPENMP=0 -DHIP=0"
OMP_DEFS="-DV100=0 -DNOACC=0 -DCUDA=0 -DOPENMP=1 -DHIP=0"
src() { for f in "$@"; do printf '%s ' "$REF/$f"; done; }
QUIET="-w"   # the reference has known %ld-into-int format warnings

# serial path, as MakefileCPU: -std=c99 -O3
$CC $CPU_DEFS -std=c99 -O3 $QUIET -fPIC -shared -Wl,-Bsymbolic -I"$REF" \
    -o "$OUT/liblapref_cpu.so" $(src simple_blas.c $SOLVER) -lm
for drv in cg_driver2:lap_cpu_cg cg_driver:lap_cpu_laplacian mm_driver:lap_cpu_spmv; do
  $CC $CPU_DEFS -std=c99 -O3 $QUIET -I"$REF" -o "$OUT/${drv#*:}" \
      $(src simple_blas.c $SOLVER ${drv%%:*}.c) -lm
done
# OpenMP path, as MakefileOMP_NOGPU: -std=c99 -O3 -fopenmp
$CC $OMP_DEFS -std=c99 -O3 -fopenmp $QUIET -fPIC -shared -Wl,-Bsymbolic -I"$REF" \
    -o "$OUT/liblapref_omp.so" $(src openmp_blas.c $SOLVER) -lm
for drv in cg_driver2:lap_ompnogpu_cg mm_driver:lap_ompnogpu_spmv; do
  $CC $OMP_DEFS -std=c99 -O3 -fopenmp $QUIET -I"$REF" -o "$OUT/${drv#*:}" \
      $(src openmp_blas.c $SOLVER ${drv%%:*}.c) -lm
done

# drop-in link: the reference's HIP-configured host code against our library
LIB=$HERE/../lap_amd
if [ -f "$LIB/liblapgfx.so" ]; then
  HIP_DEFS="-DV100=0 -DNOACC=0 -DCUDA=0 -DOPENMP=0 -DHIP=1"
  # devMem.h / hip_blas.h declarations come from the reference headers; the
  # definitions come from liblapgfx.so (no rocSPARSE / rocBLAS on the line)
  for drv in cg_driver2:lap_dropin_cg cg_driver:lap_dropin_laplacian mm_driver:lap_dropin_spmv; do
    $CC $HIP_DEFS -O3 $QUIET -I"$REF" -o "$OUT/${drv#*:}" \
        $(src $SOLVER ${drv%%:*}.c) -L"$LIB" -llapgfx -Wl,-rpath,'$ORIGIN/../../lap_amd' -lm
  done
  # the same drivers with io_utils.c left out: read_mm_file, coo_to_csr,
  # split, create_L_and_split and read_rhs come from liblapgfx.so as well
  NOIO="blas.c GS.c it_jacobi.c line_jacobi.c prec.c cg.c"
  for drv in cg_driver2:lap_dropin_io_cg cg_driver:lap_dropin_io_laplacian; do
    $CC $HIP_DEFS -O3 $QUIET -I"$REF" -o "$OUT/${drv#*:}" \
        $(src $NOIO ${drv%%:*}.c) -L"$LIB" -llapgfx -Wl,-rpath,'$ORIGIN/../../lap_amd' -lm
  done
fi
echo "build_ref: built $(ls "$OUT" | tr '\n' ' ')"
