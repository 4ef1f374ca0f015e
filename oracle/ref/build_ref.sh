#!/bin/bash
# build_ref.sh -- build the reference's Fortran hot path (OpenMp_Multicore) with the
# harness in this directory, for one set of compile-time parameters.
#
# TEST INFRASTRUCTURE ONLY: runs in the development container, where the reference
# checkout exists. The reference sources are copied to a temporary directory outside the
# repository, patched there, compiled with ROCm flang, and only the executable is written
# to oracle/_ref/ (git-ignored). No reference source is kept.
#
# Patches applied to the temporary copy (none changes the arithmetic):
#   * three dummy arguments renamed because flang rejects a dummy that shares its name
#     with a module variable (pgf90 accepted them):
#       RESZ(..., RZ) -> RZ_d   (equations.f90:723-765)
#       RESC(..., RC) -> RC_d   (equations.f90:823-867)
#       bcC(Cn)       -> Cn_d   (boundary.f90:119-159)
#   * compile-time PARAMETER values (comum.f90:29-115) overridden from the command line,
#     because the reference fixes grid size and physics constants at compile time.
#
# usage: [REF_FFLAGS=...] build_ref.sh TAG KEY=VALUE ...
#   REF_FFLAGS adds compiler flags that do not change the arithmetic (e.g. -mcmodel=medium for
#   grids whose static arrays exceed 2 GB).
#   keys: imax jmax Re Pe Sc Darcy_number Cf g Fr v_i Lhori y_up y_down Px_grid Py_grid Q_grid
#   values are Fortran literals (e.g. 1.0d-3).
set -euo pipefail
TAG=$1; shift
REF=${CFD_REFERENCE_DIR:-/root/reference}/OpenMp_Multicore
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/../_ref
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
mkdir -p "$OUT"
W=$(mktemp -d /tmp/cfdref.XXXXXX)
trap 'rm -rf "$W"' EXIT
for f in comum.f90 boundary.f90 equations.f90 nonsymetric_mesh.f90 properties_CH4.f90; do
    cp "$REF/$f" "$W/$f"
done

PX=""; PY=""; QG=""
for kv in "$@"; do
    k=${kv%%=*}; v=${kv#*=}
    case "$k" in
        imax|jmax)
            sed -i -E "s/^( *INTEGER, PARAMETER :: $k = ).*$/\1$v/" "$W/comum.f90" ;;
        Px_grid) PX=$v ;;
        Py_grid) PY=$v ;;
        Q_grid)  QG=$v ;;
        *)
            sed -i -E "s/^( *REAL\(8\), PARAMETER :: $k *= *)[^!]*/\1$v /" "$W/comum.f90" ;;
    esac
done
if [ -n "$PX$PY$QG" ]; then
    sed -i -E "s/^( *REAL\(8\), PARAMETER :: Px_grid = ).*$/\1${PX:-1.d0}, Py_grid = ${PY:-1.6d0}, Q_grid = ${QG:-1.0d0}/" "$W/comum.f90"
fi

sed -i -E '/^subroutine RESZ/,/^end subroutine RESZ/ s/\bRZ\b/RZ_d/g' "$W/equations.f90"
sed -i -E '/^subroutine RESC/,/^end subroutine RESC/ s/\bRC\b/RC_d/g' "$W/equations.f90"
sed -i -E '/^SUBROUTINE bcC/,/^END SUBROUTINE bcC/ s/\bCn\b/Cn_d/g' "$W/boundary.f90"

cp "$HERE/harness.f90" "$W/harness.f90"
( cd "$W" && "$FC" -O2 -ffp-contract=off -fopenmp ${REF_FFLAGS:-} -o "$OUT/ref_$TAG" \
      comum.f90 boundary.f90 equations.f90 nonsymetric_mesh.f90 properties_CH4.f90 harness.f90 )
grep -E "PARAMETER :: (imax|jmax|Re|Pe|Sc|Darcy_number|Cf|g|Fr|v_i|Px_grid|Lhori|y_up|y_down) " "$W/comum.f90" > "$OUT/ref_$TAG.params"
echo "built $OUT/ref_$TAG"
