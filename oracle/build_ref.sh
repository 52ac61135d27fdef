#!/bin/bash
# Build recipe for the reference solver used as a checker (never shipped, never measured as the product).
#
# Compiles the UNMODIFIED Fortran sources of the reference, where they lie, with the image's flang and
# links them against the image's MPICH (no stand-ins are written for anything).  Output goes only to
# oracle/_ref/ (git-ignored):
#   oracle/_ref/MF_LBM_multiphase.cpu   two-phase solver   (multiphase_3D/0.src/*.F90)
#   oracle/_ref/MF_LBM_singlephase.cpu  single-phase solver (singlephase_3D/0.src/*.F90)
#   oracle/_ref/wall_boundary_process   boundary-info preprocessing tool
#                                       (preprocessing/3.wall_boundary_preprocess/3d_wall_boundary_process.f90)
# The binaries run as a singleton MPI process (no mpirun needed), single thread unless OMP_NUM_THREADS
# is raised.  Flags: -O2 on generic x86-64 (no FMA contraction is possible without -march), OpenMP on.
set -euo pipefail
REF=${MFLBM_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
MPI_INC=${MPI_INC:-/opt/conda/include}
MPI_LIB=${MPI_LIB:-/opt/conda/lib}
if [ ! -d "$REF/multiphase_3D/0.src" ]; then
  echo "reference sources not found under $REF; nothing to build" >&2
  exit 3
fi
FLAGS="-cpp -O2 -fopenmp -I$MPI_INC"
build() {   # build <name> <srcdir>
  local name=$1 src=$2 obj="$OUT/obj_$1"
  mkdir -p "$obj"
  # Module.F90 defines every module the other files use, so it goes first.
  "$FC" $FLAGS -I"$src" -module-dir "$obj" -c "$src/Module.F90" -o "$obj/Module.o"
  for f in "$src"/*.F90; do
    b=$(basename "$f" .F90)
    [ "$b" = Module ] && continue
    "$FC" $FLAGS -I"$src" -module-dir "$obj" -c "$f" -o "$obj/$b.o" 2>"$obj/$b.log" || { cat "$obj/$b.log"; exit 1; }
  done
  "$FC" -fopenmp -o "$OUT/MF_LBM_$name.cpu" "$obj"/*.o -L"$MPI_LIB" -Wl,-rpath,"$MPI_LIB" -lmpifort -lmpi
}
build multiphase "$REF/multiphase_3D/0.src"
build singlephase "$REF/singlephase_3D/0.src"
mkdir -p "$OUT/obj_tool"
"$FC" -O2 -fopenmp -module-dir "$OUT/obj_tool" "$REF/preprocessing/3.wall_boundary_preprocess/3d_wall_boundary_process.f90" \
    -o "$OUT/wall_boundary_process"
echo "built: $OUT/MF_LBM_multiphase.cpu $OUT/MF_LBM_singlephase.cpu $OUT/wall_boundary_process"
