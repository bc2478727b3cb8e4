#!/usr/bin/env bash
# build_ref.sh -- compile the reference Fortran (read in place under
# /root/reference, never copied into the repository) into oracle/_ref/<tag>/.
#
# TEST INFRASTRUCTURE ONLY: the binary it produces is the parity oracle for
# the golden fixtures and the CPU baseline; the product never runs it.
#
# The reference sizes every array with the compile-time PARAMETER max_points
# and takes the flow conditions as PARAMETERs (parameter.f90:8-23), so each
# configuration gets its own build: parameter.f90 is derived into the build
# directory with these values substituted (and, when given, the WLS weight
# exponent power and the q-derivative inner_iterations, parameter.f90:21,24).
# The canonical implicit build is the alias-ordered sweep
# (implicit_forward_and_reverse_sweeps.f90) with state_update.f90; the
# alternative implicit build (implicit_backup) is the type-ordered sweep
# that also sweeps outer points (implicit_forward_and_reverse_sweeps_backup.f90)
# with state_update_implicit.f90; the explicit build links
# oracle/ref_explicit_sweep.f90 in place of the sweep module.  flang -O2 for
# the default x86-64 target, no FMA.
#
# usage: build_ref.sh <tag> <max_points> <max_iters> <mach> <aoa> <cfl>
#                     <explicit|implicit|implicit_backup> [<power> [<inner_iterations>]]
set -euo pipefail
REF=${MFCFD_REFERENCE_DIR:-/root/reference}
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
FFLAGS=${FFLAGS:-"-O2 -ffp-contract=off"}

if [ $# -lt 7 ] || [ $# -gt 9 ]; then
    echo "usage: $0 <tag> <max_points> <max_iters> <mach> <aoa> <cfl> <explicit|implicit|implicit_backup> [power [inner_iterations]]" >&2
    exit 2
fi
tag=$1; npts=$2; iters=$3; mach=$4; aoa=$5; cfl=$6; mode=$7
power=${8:-0.0}; inner=${9:-0}
[ -d "$REF" ] || { echo "reference sources not found at $REF" >&2; exit 3; }
out="$HERE/_ref/$tag"
mkdir -p "$out"

dlit() { case "$1" in *[dDeE]*) echo "$1" ;; *) echo "${1}d0" ;; esac; }
sed -e "s/^\(\s*integer, parameter :: max_points = \)26133/\1$npts/" \
    -e "s/^\(\s*integer, parameter :: max_iters = \)1000/\1$iters/" \
    -e "s/^\(\s*real\*8, parameter :: Mach = \)2.0d0/\1$(dlit "$mach")/" \
    -e "s/^\(\s*real\*8, parameter :: aoa = \)0.0d0/\1$(dlit "$aoa")/" \
    -e "s/^\(\s*real\*8, parameter :: CFL = \)5.0d0/\1$(dlit "$cfl")/" \
    -e "s/^\(\s*real\*8, parameter :: power = \)0.0d0/\1$(dlit "$power")/" \
    -e "s/^\(\s*integer, parameter :: inner_iterations = \)0/\1$inner/" \
    "$REF/parameter.f90" > "$out/parameter.f90"
grep -q "max_points = $npts" "$out/parameter.f90" || { echo "parameter substitution failed" >&2; exit 4; }
grep -q "power = $(dlit "$power")" "$out/parameter.f90" || { echo "power substitution failed" >&2; exit 4; }
grep -q "inner_iterations = $inner\$" "$out/parameter.f90" || { echo "inner_iterations substitution failed" >&2; exit 4; }

update="$REF/state_update.f90"
if [ "$mode" = explicit ]; then
    sweep="$HERE/ref_explicit_sweep.f90"
elif [ "$mode" = implicit ]; then
    sweep="$REF/implicit_forward_and_reverse_sweeps.f90"
elif [ "$mode" = implicit_backup ]; then
    sweep="$REF/implicit_forward_and_reverse_sweeps_backup.f90"
    update="$REF/state_update_implicit.f90"
else
    echo "mode must be explicit, implicit or implicit_backup" >&2; exit 2
fi

# dependency order of the canonical file set (module 'use' graph)
srcs=(
  "$out/parameter.f90" "$REF/data_structure.f90"
  "$REF/split_fluxes.f90" "$REF/octant_fluxes.f90"
  "$REF/q_variables.f90" "$REF/q_derivatives.f90"
  "$REF/qvariables_to_primitive_variables.f90" "$REF/limiters.f90"
  "$REF/compute_conserved_vector.f90" "$REF/timestep_delt.f90"
  "$REF/interior_flux_dGxpos.f90" "$REF/interior_flux_dGxneg.f90"
  "$REF/interior_flux_dGypos.f90" "$REF/interior_flux_dGyneg.f90"
  "$REF/interior_flux_dGzpos.f90" "$REF/interior_flux_dGzneg.f90"
  "$REF/wall_flux_dGxpos.f90" "$REF/wall_flux_dGxneg.f90"
  "$REF/wall_flux_dGypos.f90" "$REF/wall_flux_dGyneg.f90" "$REF/wall_flux_dGzneg.f90"
  "$REF/outer_flux_dGxpos.f90" "$REF/outer_flux_dGxneg.f90"
  "$REF/outer_flux_dGypos.f90" "$REF/outer_flux_dGyneg.f90" "$REF/outer_flux_dGzpos.f90"
  "$REF/flux_residual.f90" "$REF/get_primitive_vector.f90"
  "$REF/implicit_D_inv_interior.f90" "$REF/implicit_D_inv_wall.f90" "$REF/implicit_D_inv_outer.f90"
  "$REF/implicit_LdelU_forward_sweep_interior.f90" "$REF/implicit_LdelU_forward_sweep_wall.f90"
  "$REF/implicit_LdelU_forward_sweep_outer.f90"
  "$REF/implicit_UdelU_backward_sweep_interior.f90" "$REF/implicit_UdelU_backward_sweep_wall.f90"
  "$REF/implicit_UdelU_backward_sweep_outer.f90"
  "$sweep"
  "$REF/primitive_to_conserved_vector.f90" "$REF/conserved_to_primitive_vector.f90"
  "$REF/conserved_vector_Ubar.f90" "$update" "$REF/fpi_solver.f90"
  "$REF/generate_connectivity.f90" "$REF/implicit_aliasing.f90"
  "$REF/point_preprocessor.f90" "$REF/initial_conditions.f90"
  "$HERE/ref_driver.f90"
)
objs=()
for s in "${srcs[@]}"; do
    o="$out/$(basename "${s%.f90}").o"
    $FC $FFLAGS -module-dir "$out" -I "$out" -c "$s" -o "$o"
    objs+=("$o")
done
$FC $FFLAGS "${objs[@]}" -o "$out/ref_driver"
echo "$out/ref_driver"
