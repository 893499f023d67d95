#!/bin/sh
# ORACLE — test infrastructure only.
#
# Streams the reference's Solver.cpp with the edits of one variant into a
# compiler command that reads the translation unit from stdin (oracle/Makefile
# passes g++ ... -x c++ -).  Usage: ref_variant.sh VARIANT REFERENCE CMD...
#
#   visc   Solver.cpp:792 `State TotalFlux = FluxEuler;` becomes the viscous
#          form of 787, `FluxEuler - FluxVisc`; the centroid offset of
#          729-730 is 0 on a boundary face (f.neighbor = -1), where the
#          reference reads mesh.cells[-1]
#   plate  visc + the active getBoundaryGhostState (806-863) removed and the
#          laminar flat plate / SWBLI one (1088-1139) uncommented
#   super  the active getBoundaryGhostState (806-863) removed and the NASA
#          supersonic one (973-1011) uncommented
#
# The source has CRLF line ends; they are dropped first.  Every line an edit
# addresses is checked against its expected text before anything is compiled,
# so a different revision of the reference fails here instead of compiling
# into something else.  The exit status is the compiler's.
set -eu
V="$1"
SRC="$2/src/Solver.cpp"
shift 2
TXT=$(tr -d '\r' < "$SRC")

expect() {  # line text
  got=$(printf '%s\n' "$TXT" | sed -n "$1p")
  if [ "$got" != "$2" ]; then
    echo "ref_variant.sh: $SRC:$1 reads '$got', expected '$2'" >&2
    exit 1
  fi
}

VISC='729s/mesh\.cells\[f\.neighbor\]\.x_c - mesh\.cells\[f\.owner\]\.x_c/(f.neighbor >= 0 ? mesh.cells[f.neighbor].x_c - mesh.cells[f.owner].x_c : 0.0)/
730s/mesh\.cells\[f\.neighbor\]\.y_c - mesh\.cells\[f\.owner\]\.y_c/(f.neighbor >= 0 ? mesh.cells[f.neighbor].y_c - mesh.cells[f.owner].y_c : 0.0)/
792s/State TotalFlux = FluxEuler;/State TotalFlux = FluxEuler - FluxVisc;/'
GHOST='State Solver::getBoundaryGhostState(const State& U_in, const Face& f) {'

check_visc() {
  expect 729 '            double dx_LR = mesh.cells[f.neighbor].x_c - mesh.cells[f.owner].x_c;'
  expect 730 '            double dy_LR = mesh.cells[f.neighbor].y_c - mesh.cells[f.owner].y_c;'
  expect 787 '        // State TotalFlux = FluxEuler - FluxVisc;'
  expect 792 '        State TotalFlux = FluxEuler;'
}
check_bump_ghost() {
  expect 806 "$GHOST"
  expect 863 '}'
}

case "$V" in
  visc)
    check_visc
    printf '%s\n' "$TXT" | sed -e "$VISC" | "$@"
    ;;
  plate)
    check_visc
    check_bump_ghost
    expect 1087 '// for laminar flat plate / SWBLI mach 5 case'
    expect 1088 "// $GHOST"
    expect 1139 '// }'
    printf '%s\n' "$TXT" | sed -e "$VISC" -e '806,863d' -e '1088,1139s/^\/\/ \{0,1\}//' | "$@"
    ;;
  super)
    check_bump_ghost
    expect 971 '// for NASA supersonic test case'
    expect 973 "// $GHOST"
    expect 1011 '// }'
    printf '%s\n' "$TXT" | sed -e '806,863d' -e '973,1011s/^\/\/ \{0,1\}//' | "$@"
    ;;
  *)
    echo "ref_variant.sh: unknown variant '$V' (visc, plate, super)" >&2
    exit 1
    ;;
esac
