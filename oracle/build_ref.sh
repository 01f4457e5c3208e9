#!/usr/bin/env bash
# Build the reference hot-path modules into oracle/_ref/ -- TEST INFRASTRUCTURE.
#
# Compiles the UNMODIFIED reference modules mhd_field_ops, mhd_change_states,
# mhd_derivatives, mhd_bc and mhd_flux from $CHIMERA_REF_SRC (default
# /root/reference/src) with amdflang, together with oracle/ref_harness.f90, into
#   oracle/_ref/libchimera_ref_s<solver>_l<limiter>.so
# for riemann_solver in {1 Rusanov, 2 HLLE, 3 HLLD} x slope_limiter in {1 Van Leer,
# 2 MC}.  Those two are compile-time `parameter`s of the reference's
# mhd_config.f90 (lines 30 and 36), so each variant compiles a copy of
# mhd_config.f90 with that one value substituted; the copy lives in a temporary
# directory that is removed afterwards and never enters the repository.
#
# Flags: -O3 -fopenmp -ffp-contract=off (no FMA contraction, so the bits are
# comparable with the C oracle and the HIP kernels).
#
# When chimera_amd/libchimera_hip.so exists, each variant is also linked a second
# time with the product's Fortran drop-in modules (chimera_amd/fortran/mhd_*.f90)
# in place of the five reference modules:
#   oracle/_ref/libchimera_dropin_s<solver>_l<limiter>.so
# -- the same harness and the reference's mhd_config, every routine on the GPU.
# With chimera_amd/libchimera_h5.so also present, the reference's main program is
# linked the same way (see the end of this script).
set -euo pipefail

HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
SRC="${CHIMERA_REF_SRC:-/root/reference/src}"
OUT="$HERE/_ref"
SHIMS="$HERE/../chimera_amd/fortran"
LIBHIP="$HERE/../chimera_amd/libchimera_hip.so"
FC="${FC:-amdflang}"
FFLAGS="-O3 -fPIC -fopenmp -ffp-contract=off"

if [ ! -f "$SRC/mhd_config.f90" ]; then
    echo "build_ref.sh: reference sources not found at $SRC" >&2
    exit 3
fi
command -v "$FC" >/dev/null || { echo "build_ref.sh: $FC not found" >&2; exit 3; }
mkdir -p "$OUT"

TMP="$(mktemp -d /tmp/chimera_ref.XXXXXX)"
trap 'rm -rf "$TMP"' EXIT

solvers=(x RIEMANN_RUSANOV RIEMANN_HLLE RIEMANN_HLLD)
limiters=(x LIMITER_VAN_LEER LIMITER_MC)

for s in 1 2 3; do
  for l in 1 2; do
    B="$TMP/s${s}_l${l}"
    mkdir -p "$B"
    sed -E \
      -e "s/^([[:space:]]*integer, parameter :: slope_limiter[[:space:]]*=[[:space:]]*)LIMITER_[A-Z_]+/\1${limiters[$l]}/" \
      -e "s/^([[:space:]]*integer, parameter :: riemann_solver[[:space:]]*=[[:space:]]*)RIEMANN_[A-Z]+/\1${solvers[$s]}/" \
      "$SRC/mhd_config.f90" > "$B/mhd_config.f90"
    grep -Eq "slope_limiter[[:space:]]*=[[:space:]]*${limiters[$l]}" "$B/mhd_config.f90"
    grep -Eq "riemann_solver[[:space:]]*=[[:space:]]*${solvers[$s]}" "$B/mhd_config.f90"
    objs=()
    for m in "$B/mhd_config.f90" "$SRC/mhd_field_ops.f90" "$SRC/mhd_change_states.f90" \
             "$SRC/mhd_derivatives.f90" "$SRC/mhd_bc.f90" "$SRC/mhd_flux.f90" \
             "$HERE/ref_harness.f90"; do
      o="$B/$(basename "$m" .f90).o"
      $FC $FFLAGS -module-dir "$B" -I "$B" -c "$m" -o "$o"
      objs+=("$o")
    done
    $FC -shared -fopenmp -o "$OUT/libchimera_ref_s${s}_l${l}.so" "${objs[@]}"

    if [ -f "$LIBHIP" ]; then
      D="$B/dropin"
      mkdir -p "$D"
      dobjs=()
      for m in "$B/mhd_config.f90" "$SHIMS/chimera_c.f90" "$SHIMS/mhd_field_ops.f90" \
               "$SHIMS/mhd_change_states.f90" "$SHIMS/mhd_derivatives.f90" "$SHIMS/mhd_bc.f90" \
               "$SHIMS/mhd_flux.f90" "$HERE/ref_harness.f90"; do
        o="$D/$(basename "$m" .f90).o"
        $FC $FFLAGS -module-dir "$D" -I "$D" -c "$m" -o "$o"
        dobjs+=("$o")
      done
      $FC -shared -fopenmp -o "$OUT/libchimera_dropin_s${s}_l${l}.so" "${dobjs[@]}" \
          -L"$(dirname "$LIBHIP")" -lchimera_hip -Wl,-rpath,'$ORIGIN/../../chimera_amd'
    fi
  done
done

# The reference's own program (main.f90, mhd_init.f90 and its mhd_config, unmodified):
#   oracle/_ref/chimera_ref_main     the reference modules, with only mhd_write_h5
#                                    swapped for the drop-in writer (the reference's
#                                    needs the gfortran HDF5 module); runs on the CPU
#   oracle/_ref/chimera_dropin_main  the drop-in modules and writer: every hot-path
#                                    routine on the GPU
# both with the reference CLI: <problem_type> <N> <seed> [output_path] [h5_filename] ...
LIBH5="$HERE/../chimera_amd/libchimera_h5.so"
if [ -f "$LIBH5" ]; then
  R="$TMP/refmain"
  mkdir -p "$R"
  robjs=()
  for m in "$SRC/mhd_config.f90" "$SRC/mhd_field_ops.f90" "$SRC/mhd_change_states.f90" \
           "$SRC/mhd_derivatives.f90" "$SRC/mhd_bc.f90" "$SRC/mhd_flux.f90" \
           "$SHIMS/mhd_write_h5.f90" "$SRC/mhd_init.f90" "$SRC/main.f90"; do
    o="$R/$(basename "$m" .f90).o"
    $FC $FFLAGS -module-dir "$R" -I "$R" -c "$m" -o "$o"
    robjs+=("$o")
  done
  $FC -fopenmp -o "$OUT/chimera_ref_main" "${robjs[@]}" \
      -L"$(dirname "$LIBH5")" -lchimera_h5 -Wl,-rpath,'$ORIGIN/../../chimera_amd'
fi
if [ -f "$LIBHIP" ] && [ -f "$LIBH5" ]; then
  M="$TMP/main"
  mkdir -p "$M"
  mobjs=()
  for m in "$SRC/mhd_config.f90" "$SHIMS/chimera_c.f90" "$SHIMS/mhd_field_ops.f90" \
           "$SHIMS/mhd_change_states.f90" "$SHIMS/mhd_derivatives.f90" "$SHIMS/mhd_bc.f90" \
           "$SHIMS/mhd_flux.f90" "$SHIMS/mhd_write_h5.f90" "$SRC/mhd_init.f90" "$SRC/main.f90"; do
    o="$M/$(basename "$m" .f90).o"
    $FC $FFLAGS -module-dir "$M" -I "$M" -c "$m" -o "$o"
    mobjs+=("$o")
  done
  $FC -fopenmp -o "$OUT/chimera_dropin_main" "${mobjs[@]}" \
      -L"$(dirname "$LIBHIP")" -lchimera_hip -lchimera_h5 -Wl,-rpath,'$ORIGIN/../../chimera_amd'
fi
echo "build_ref.sh: built $(ls "$OUT"/libchimera_ref_*.so | wc -l) reference and" \
     "$(ls "$OUT"/libchimera_dropin_*.so 2>/dev/null | wc -l) drop-in variants into $OUT" \
     "$( [ -f "$OUT/chimera_ref_main" ] && echo '+ chimera_ref_main')" \
     "$( [ -f "$OUT/chimera_dropin_main" ] && echo '+ chimera_dropin_main')"
