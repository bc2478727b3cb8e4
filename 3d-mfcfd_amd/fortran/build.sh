#!/usr/bin/env bash
# build.sh -- link the reference's own host program against the GPU drop-in.
#
# Compiles the reference's parameter_mod, data_structure_mod, host-side
# routines and main program (meshfree_solver.f90, q_lskum.f90,
# point_preprocessor.f90, initial_conditions.f90, generate_connectivity.f90,
# implicit_aliasing.f90; read in place from the reference directory)
# together with mfcfd_hip_binding.f90 and mfcfd_dropin.f90, which replace
# every hot-path module, and mfcfd_postprocess.f90, which replaces
# post_processor_mod (the reference's version stops on a Fortran unit
# conflict under flang), and links libmfcfd_hip.so.  The program is the
# reference's unchanged driver running the HIP hot path.  Because it contains compiled reference code, the
# output goes to oracle/_ref/dropin_<tag>/ with the other reference builds.
#
# usage: build.sh [<tag> <max_points> <max_iters> <mach> <aoa> <cfl>]
#   with no arguments: the golden-fixture configurations
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
PKG="$(dirname "$HERE")"
ROOT="$(dirname "$PKG")"
REF=${MFCFD_REFERENCE_DIR:-/root/reference}
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
FFLAGS=${FFLAGS:-"-O2"}

if [ ! -d "$REF" ]; then
    echo "reference sources not found at $REF; nothing to link the drop-in against" >&2
    exit 0
fi
[ -f "$PKG/libmfcfd_hip.so" ] || { echo "build libmfcfd_hip.so first" >&2; exit 3; }

one() {
    local tag=$1 npts=$2 iters=$3 mach=$4 aoa=$5 cfl=$6
    local out="$ROOT/oracle/_ref/dropin_$tag"
    mkdir -p "$out"
    dlit() { case "$1" in *[dDeE]*) echo "$1" ;; *) echo "${1}d0" ;; esac; }
    sed -e "s/^\(\s*integer, parameter :: max_points = \)26133/\1$npts/" \
        -e "s/^\(\s*integer, parameter :: max_iters = \)1000/\1$iters/" \
        -e "s/^\(\s*real\*8, parameter :: Mach = \)2.0d0/\1$(dlit "$mach")/" \
        -e "s/^\(\s*real\*8, parameter :: aoa = \)0.0d0/\1$(dlit "$aoa")/" \
        -e "s/^\(\s*real\*8, parameter :: CFL = \)5.0d0/\1$(dlit "$cfl")/" \
        "$REF/parameter.f90" > "$out/parameter.f90"
    grep -q "max_points = $npts" "$out/parameter.f90" || { echo "parameter substitution failed" >&2; exit 4; }
    local srcs=(
        "$out/parameter.f90" "$REF/data_structure.f90"
        "$HERE/mfcfd_hip_binding.f90" "$HERE/mfcfd_dropin.f90"
        "$REF/generate_connectivity.f90" "$REF/implicit_aliasing.f90"
        "$REF/point_preprocessor.f90" "$REF/initial_conditions.f90" "$HERE/mfcfd_postprocess.f90"
        "$REF/q_lskum.f90" "$REF/meshfree_solver.f90"
    )
    local objs=()
    for s in "${srcs[@]}"; do
        local o="$out/$(basename "${s%.f90}").o"
        $FC $FFLAGS -module-dir "$out" -I "$out" -c "$s" -o "$o"
        objs+=("$o")
    done
    $FC $FFLAGS "${objs[@]}" -L"$PKG" -lmfcfd_hip -Wl,-rpath,'$ORIGIN/../../../3d-mfcfd_amd' \
        -o "$out/meshfree_solver"
    echo "$out/meshfree_solver"
}

if [ $# -eq 6 ]; then
    one "$@"
else
    one explicit_blocks 1059 40 0.6 2.0 0.2
    one implicit_morton 1819 30 0.84 3.06 2.0
fi
