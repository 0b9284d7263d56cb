#!/usr/bin/env bash
# Builds the reference's own CPU solver core from its sources under
# $REF/src into oracle/_ref/ref_<scenario>, one binary per compile-time
# parameter set, driven by oracle/ref_harness.cpp. TEST INFRASTRUCTURE ONLY.
#
# Scenario parameter sets are either the reference's own test shadows
# ($REF/tests/params/<dir>/params.h, the mechanism of tests/CMakeLists.txt:3-9)
# or $REF/src/params.h with grid size / LAMBDA / ALPHA / thread count edited by
# sed, plus $REF/src/sim_config.h with the advection scheme edited. Nothing
# from the reference is copied into the repository: generated headers and
# binaries land only in oracle/_ref/ (git-ignored).
#
# Compile flags follow the reference's Release CPU build (-O3 -DNDEBUG
# -fopenmp, no -march; CMakeLists.txt:50-62).
set -euo pipefail
REF=${REF:-/root/reference}
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/_ref"
CXX=${CXX:-g++}
if [ ! -f "$REF/src/lbm_solver.tpp" ]; then
    echo "build_ref: no reference sources at $REF; skipping" >&2
    exit 0
fi
mkdir -p "$OUT"

# name | params source (test:<dir> or gen:nx,ny,nz,lambda,alpha,threads) | advection
SCENARIOS=(
  "unit|test:unit|Centred"
  "coupled|test:coupled|Centred"
  "bccheck|test:bc_check|Centred"
  "qtrelax|test:qtrelax|Centred"
  "poiseuille|test:poiseuille|Centred"
  "act|gen:12,10,8,0.3,0.023,1|Centred"
  "actup|gen:12,10,8,0.7,0.046,1|Upwind"
  "c0|gen:32,32,32,0.3,0.023,1|Centred"
  "c1|gen:20,100,100,0.5,0.023,8|Centred"
  "c2|gen:256,256,256,0.7,0.046,16|Upwind"
)
ONLY=${1:-}

build_one() {
    local spec="$1"
    IFS='|' read -r name psrc adv <<< "$spec"
    if [ -n "$ONLY" ] && [[ " $ONLY " != *" $name "* ]]; then return 0; fi
    local cfg="$OUT/cfg_$name"
    mkdir -p "$cfg"
    local pinc
    if [[ "$psrc" == test:* ]]; then
        pinc="$REF/tests/params/${psrc#test:}"
    else
        IFS=',' read -r gnx gny gnz glam galpha gthr <<< "${psrc#gen:}"
        sed -e "s/\(inline constexpr int nx = \)[0-9]*/\1$gnx/" \
            -e "s/\(inline constexpr int ny = \)[0-9]*/\1$gny/" \
            -e "s/\(inline constexpr int nz = \)[0-9]*/\1$gnz/" \
            -e "s/\(inline constexpr int kNumOMPThreads = \)[0-9]*/\1$gthr/" \
            -e "s/\(inline constexpr double LAMBDA = \)[0-9.]*/\1$glam/" \
            -e "s/\(inline constexpr double ALPHA = \)[0-9.]*/\1$galpha/" \
            "$REF/src/params.h" > "$cfg/params.h"
        pinc="$cfg"
    fi
    sed -e "s/\(inline constexpr Advection kQAdvection = Advection::\)[A-Za-z]*/\1$adv/" \
        "$REF/src/sim_config.h" > "$cfg/sim_config.h"
    # sim_config.h goes first; params.h from the scenario directory next.
    "$CXX" -std=c++20 -O3 -DNDEBUG -fopenmp -I"$cfg" -I"$pinc" -I"$REF/src" \
        "$HERE/ref_harness.cpp" "$REF/src/fluid_fields.cc" "$REF/src/qtensor_fields.cc" \
        "$REF/src/analysis_fields.cc" -o "$OUT/ref_$name"
    echo "built $OUT/ref_$name"
}

pids=()
for s in "${SCENARIOS[@]}"; do
    build_one "$s" &
    pids+=($!)
    if [ ${#pids[@]} -ge 4 ]; then wait "${pids[0]}"; pids=("${pids[@]:1}"); fi
done
for p in "${pids[@]}"; do wait "$p"; done
