#!/usr/bin/env bash
# build_ref.sh -- compile the reference's own hot-path Fortran modules, read in
# place under /root/reference, with flang and the project's single-domain
# driver (oracle/ref_driver.f90), for ONE compile-time configuration, and run
# it once.  Everything it writes goes under oracle/_ref/<name>/.
#
# The reference selects precision and every parameter at compile time
# (source/settings.f90:7-93, CMakeLists.txt:97-108), so each configuration is
# its own build, exactly like the reference's FFB_SETTINGS_DEFINITIONS jobs.
#
# usage: oracle/build_ref.sh <name> <fp64|fp32> FFB_SIM_MODE=<m> FFB_N_X=<nx> ...
#   Any FFB_* definition not given takes the default below.
#   REF_OPT (default "-O0") sets the optimisation flags; REF_RUN=0 builds
#   without running; REF_KEEP=0 removes everything but state.bin after the run.
# output: oracle/_ref/<name>/ref_driver and oracle/_ref/<name>/state.bin
#
# Test infrastructure only; needs the reference tree, so it runs in the build
# container, never on a GPU host.
set -euo pipefail

REF=${FFB_REFERENCE_ROOT:-/root/reference}
HERE=$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)
FLANG=${FLANG:-flang}

name=$1; prec=$2; shift 2
out="$HERE/_ref/$name"
mkdir -p "$out"

declare -A defs=(
  [FFB_SIM_MODE]=4 [FFB_N_X]=16 [FFB_N_Y]=16 [FFB_N_STEPS]=10
  [FFB_RHO_0]=1.0_FP [FFB_OMEGA]=1.1_FP [FFB_U_MAX]=0.1_FP [FFB_N_SIN]=2.0_FP
  [FFB_U_WALL]=0.1_FP [FFB_U_LID]=0.1_FP [FFB_RHO_IN]=1.00005_FP [FFB_RHO_OUT]=0.99995_FP
  [FFB_USE_UNROLLED_KERNELS]=.true. [FFB_EXPORT_MACROS]=.false.
  [FFB_EXPORT_ENDPOINT_STATES]=.true. [FFB_EXPORT_INTERVAL]=10000
  [FFB_INTERACTIVE_PROGRESS]=.false. [FFB_PROGRESS_INTERVAL]=1
)
for kv in "$@"; do defs[${kv%%=*}]=${kv#*=}; done

read -r -a opt <<< "${REF_OPT:--O0}"
flags=(-cpp "${opt[@]}" -ffp-contract=off -DFFB_USE_CMAKE_SETTINGS '-DFFB_EXPORT_NUM="ref"')
[ "$prec" = fp64 ] && flags+=(-DFFB_FP64)
for k in "${!defs[@]}"; do flags+=("-D$k=${defs[$k]}"); done

src=(
  "$REF/source/settings.f90"
  "$REF/source/domain.f90"
  "$REF/source/layouts/xyd/exchange.f90"
  "$REF/source/layouts/xyd/modes/shear_wave.f90"
  "$REF/source/layouts/xyd/modes/couette_flow.f90"
  "$REF/source/layouts/xyd/modes/sliding_lid.f90"
  "$REF/source/layouts/xyd/modes/poiseuille_flow.f90"
  "$REF/source/layouts/xyd/initialization.f90"
  "$REF/source/layouts/xyd/simulation.f90"
)
objs=()
cd "$out"
for s in "${src[@]}"; do
  b=$(basename "$s" .f90)
  case "$b" in
    domain|exchange)
      # coarray modules: flang writes their .mod (types only) and then stops
      # at lowering; no object is needed because the hot-path modules use
      # only their derived types.
      "$FLANG" "${flags[@]}" -fsyntax-only "$s" 2>/dev/null || true
      [ -f "$b.mod" ] || { echo "build_ref: no $b.mod from $s" >&2; exit 1; }
      ;;
    *)
      "$FLANG" "${flags[@]}" -c "$s" -o "$b.o" 2> "$b.log" || { cat "$b.log" >&2; exit 1; }
      objs+=("$b.o")
      ;;
  esac
done
"$FLANG" "${flags[@]}" -c "$HERE/ref_driver.f90" -o ref_driver.o
"$FLANG" -o ref_driver ref_driver.o "${objs[@]}"
rm -f ./*.o ./*.mod ./*.log
if [ "${REF_RUN:-1}" = 1 ]; then
  ./ref_driver "$out/state.bin"
  echo "build_ref: $name -> $out/state.bin"
  [ "${REF_KEEP:-1}" = 1 ] || rm -f ref_driver
else
  echo "build_ref: $name -> $out/ref_driver"
fi
