#!/usr/bin/env bash
# Builds the reference's own single-component path into oracle/_ref/ (test
# infrastructure only; git-ignored, but it travels with the tree so the CPU
# baseline can be timed on the GPU host).
#
# The sources are compiled where they lie under /root/reference/src and are
# never copied.  Configuration that the reference takes from its editable
# defines.hpp is given as -D flags instead (the header's guard DEFINES_H is
# pre-defined so the shipped case selection, a MULTICOMP flame, is not
# applied).  The single-component path uses no Cantera symbol, so the guard
# CANTERA_H of cantera.hpp is pre-defined as well; no header, library or code
# stands in for anything.
#
# Outputs:
#   oracle/_ref/ref_driver   -O2, no fast-math: golden vectors
#   oracle/_ref/ref_timer    the Makefile's -O3 -ffast-math -fopenmp flags
#                            (-march=x86-64-v3 instead of -march=native so the
#                            binary runs on the GPU host's CPU): CPU baseline
#   oracle/_ref/ref_driver_conj   as ref_driver with -DCONJUGATE (the
#                            conjugate heat-transfer variant of defines.hpp)
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${LBM_REFERENCE_DIR:-/root/reference}"
SRC="$REF/src"
OUT="$HERE/_ref"
if [ ! -d "$SRC" ]; then
    echo "build_ref: $SRC not present; skipping" >&2
    exit 0
fi
mkdir -p "$OUT/obj_golden" "$OUT/obj_timer" "$OUT/obj_conj"
CXX="${CXX:-g++}"
DEFS="-DDEFINES_H -DCANTERA_H -DD3Q27 -DNDIM=3 -DPARALLEL -DNUM_THREADS=8 -DSPECIES_MIN=1E-8
      -DTYPE_F=0 -DTYPE_S=1 -DTYPE_P=2 -DTYPE_A=3 -DTYPE_FS=4 -DTYPE_I=5 -DTYPE_O=6 -DTYPE_I_C=7
      -DTYPE_O_C=8 -DTYPE_Q=9 -DVEL0=0.1 -DRHO0=1.0 -DTEMP0=0.025 -DLIMITER_TYPE=limiterVanleer
      -DTAYLOR_GREEN_3D"
UNITS="lbm-main lbm-collision lbm-streaming lbm-bc lbm-moment lbm-edf lbm-init FD math-util output restart_file"

build_one() {  # $1 = obj dir, $2 = flags, $3 = output binary
    local objs=()
    for u in $UNITS; do
        "$CXX" -std=gnu++17 $2 -fopenmp $DEFS -I"$SRC/headers" -c "$SRC/$u.cpp" -o "$1/$u.o"
        objs+=("$1/$u.o")
    done
    "$CXX" -std=gnu++17 $2 -fopenmp $DEFS -I"$SRC/headers" "$HERE/ref_driver.cpp" "${objs[@]}" -o "$3"
}

build_one "$OUT/obj_golden" "-O2" "$OUT/ref_driver"
build_one "$OUT/obj_timer" "-O3 -ffast-math -march=x86-64-v3" "$OUT/ref_timer"
build_one "$OUT/obj_conj" "-O2 -DCONJUGATE" "$OUT/ref_driver_conj"
rm -rf "$OUT/obj_golden" "$OUT/obj_timer" "$OUT/obj_conj"
echo "build_ref: built $OUT/ref_driver, $OUT/ref_timer and $OUT/ref_driver_conj"
