#!/usr/bin/env bash
# Build the compiled-reference checker into oracle/_ref/ (test infrastructure
# only).  The reference's C++ sources are compiled where they lie under
# $REF; nothing is copied into the repository.  The reference configures
# nghosts and maxlevs as compile-time constants in defs.hpp (defs.hpp:27-28),
# so each variant gets a constants-only patched defs.hpp in a temporary
# directory outside the repository; `-I<tmp> -I-` makes the reference's
# `#include "defs.hpp"` resolve to it.  The temporary headers are deleted at
# the end, so no reference source ever sits in the tree or travels with it.
#
# Variants: ng<G>_lv<L>      serial, -O3 -ffp-contract=off (the oracle build)
#           ng<G>_lv<L>_omp  -fopenmp -DOPENMP, generic x86-64, no FMA contraction
#                            (CPU baseline with the oracle's arithmetic)
#           ng<G>_lv<L>_omp_native  the reference makefile's own flags
#                            (makefile:9-19: -O3 -march -funroll-loops -fopenmp,
#                            FMAs contracted) with -march=x86-64-v3 in place of
#                            -march=native: built here, it runs on other hosts
#                            (timing only: its FMAs change the results)
#           ng<G>_lv<L>_ct2  serial, CTtype = 2 (defs.hpp:34, the upwind corner
#                            EMF of CT2D.cpp:28-75) instead of 1
#           ng<G>_lv<L>_bs   serial; its "rotor" is the Balsara-Spicer rotor of
#                            the rotor workload's text (p = 1, v0 = 2,
#                            Bx = 5/sqrt(4 pi)): a temporary copy of
#                            IC2Dtype.cpp with only those constants changed
# Output: oracle/_ref/<variant>/refrun  (harness + reference objects)
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${REF:-/root/reference}"
OUT="$HERE/_ref"
JOBS="${JOBS:-8}"
VARIANTS="${VARIANTS:-ng2_lv1 ng2_lv3 ng2_lv4 ng2_lv5 ng2_lv6 ng4_lv3 ng4_lv5 ng2_lv4_omp ng2_lv5_omp ng2_lv5_omp_native ng2_lv5_bs ng2_lv4_ct2 ng4_lv3_ct2}"

if [ ! -f "$REF/defs.hpp" ]; then
  echo "build_ref: reference not present at $REF; skipping" >&2
  exit 0
fi

TMP="$(mktemp -d /tmp/amr2d_refinc.XXXXXX)"
trap 'rm -rf "$TMP"' EXIT

# renames that route main.cpp's phase calls into the harness wrappers
rename_flags() {
  case "$1" in
    main.cpp)          echo "-Dmain=ref_main" ;;
    boundary.cpp)      echo "-Dboundary=ref_boundary" ;;
    boundaryfc.cpp)    echo "-Dboundaryfc=ref_boundaryfc" ;;
    boundaryec.cpp)    echo "-Dboundaryec=ref_boundaryec" ;;
    CT2D.cpp)          echo "-DCT2D=ref_CT2D" ;;
    MUSCL2D.cpp)       echo "-DMUSCL2D=ref_MUSCL2D" ;;
    admesh.cpp)        echo "-Dadmesh=ref_admesh" ;;
    locate_bounds.cpp) echo "-Dlocate_bounds=ref_locate_bounds" ;;
    savedata.cpp)      echo "-Dsave4py=ref_save4py" ;;
    criteria.cpp)      echo "-Dcriteria=ref_criteria" ;;
    *) echo "" ;;
  esac
}

build_variant() {
  local v="$1"
  local ng lv omp
  ng="$(echo "$v" | sed -E 's/^ng([0-9]+)_.*/\1/')"
  lv="$(echo "$v" | sed -E 's/^ng[0-9]+_lv([0-9]+).*/\1/')"
  omp=0; [[ "$v" == *_omp ]] && omp=1
  native=0; [[ "$v" == *_omp_native ]] && native=1
  bs=0; [[ "$v" == *_bs ]] && bs=1
  ct2=0; [[ "$v" == *_ct2 ]] && ct2=1
  local inc="$TMP/$v" obj="$OUT/$v/obj" bin="$OUT/$v/refrun"
  mkdir -p "$inc" "$obj"
  sed -e "s/^const int nghosts=2;/const int nghosts=${ng};/" \
      -e "s/^const int maxlevs=5;/const int maxlevs=${lv};/" "$REF/defs.hpp" > "$inc/defs.hpp"
  grep -q "^const int nghosts=${ng};" "$inc/defs.hpp" || { echo "patch nghosts failed" >&2; return 1; }
  grep -q "^const int maxlevs=${lv};" "$inc/defs.hpp" || { echo "patch maxlevs failed" >&2; return 1; }
  if [ "$ct2" = 1 ]; then
    sed -i -e "s/^const int CTtype=1;/const int CTtype=2;/" "$inc/defs.hpp"
    grep -q "^const int CTtype=2;" "$inc/defs.hpp" || { echo "patch CTtype failed" >&2; return 1; }
  fi
  if [ "$bs" = 1 ]; then
    # rotor constants of IC2Dtype.cpp:79-80, 100-101, 104-105 (v0 = 2: both
    # velocity components doubled, an exact scaling)
    sed -e 's|pinit\[0\]=0\.5;|pinit[0]=1.0;|' \
        -e 's|Bxinit\[0\]=2\.5/(sqrt(4\*pi));|Bxinit[0]=5.0/(sqrt(4*pi));|' \
        -e 's|\]=-(y-0\.5)/0\.1;|]=-2.*(y-0.5)/0.1;|' \
        -e 's|\]=(x-0\.5)/0\.1;|]=2.*(x-0.5)/0.1;|' \
        -e 's|\]=(-(y-0\.5)\*fr/0\.1);|]=(-2.*(y-0.5)*fr/0.1);|' \
        -e 's|\]=(x-0\.5)\*fr/0\.1;|]=2.*(x-0.5)*fr/0.1;|' "$REF/IC2Dtype.cpp" > "$inc/IC2Dtype.cpp"
    for pat in 'pinit\[0\]=1\.0;' 'Bxinit\[0\]=5\.0/' '=-2\.\*(y-0\.5)/0\.1;' '=2\.\*(x-0\.5)/0\.1;' \
               '=(-2\.\*(y-0\.5)\*fr/0\.1);' '=2\.\*(x-0\.5)\*fr/0\.1;'; do
      [ "$(grep -c "$pat" "$inc/IC2Dtype.cpp")" = 1 ] || { echo "patch rotor_bs failed: $pat" >&2; return 1; }
    done
  fi
  local flags="-O3 -std=c++17 -ffp-contract=off -DDOUBLE_PREC -w"
  if [ "$omp" = 1 ]; then flags="-O3 -std=c++17 -funroll-loops -ffp-contract=off -DDOUBLE_PREC -DOPENMP -fopenmp -w"; fi
  if [ "$native" = 1 ]; then flags="-O3 -std=c++17 -march=x86-64-v3 -funroll-loops -DDOUBLE_PREC -DOPENMP -fopenmp -w"; fi
  # skip when up to date
  if [ -x "$bin" ] && [ "$bin" -nt "$HERE/refharness/harness.cpp" ] && [ "$bin" -nt "$HERE/build_ref.sh" ] \
     && [ -z "$(find "$REF" -newer "$bin" -name '*.[ch]pp' -print -quit)" ]; then
    return 0
  fi
  local objs=()
  for src in "$REF"/*.cpp; do
    local b; b="$(basename "$src" .cpp)"
    [ -f "$inc/$b.cpp" ] && src="$inc/$b.cpp"
    # shellcheck disable=SC2046
    g++ $flags $(rename_flags "$b.cpp") -I"$inc" -I- -I"$REF" -c "$src" -o "$obj/$b.o" 2>"$obj/$b.err" || { cat "$obj/$b.err" >&2; return 1; }
    objs+=("$obj/$b.o")
  done
  g++ $flags -I"$inc" -I- -I"$REF" -c "$HERE/refharness/harness.cpp" -o "$obj/harness.o" 2>"$obj/harness.err" || { cat "$obj/harness.err" >&2; return 1; }
  g++ $flags "${objs[@]}" "$obj/harness.o" -Wl,--wrap=system -o "$bin"
  echo "built $bin"
}

export -f build_variant rename_flags
export REF OUT TMP HERE
mkdir -p "$OUT"
printf '%s\n' $VARIANTS | xargs -P "$JOBS" -I{} bash -c 'build_variant {} || exit 255'
