#!/bin/bash
# Builds oracle/_ref/libvpref.so: the reference's own C++ cubature cores,
# compiled from where they lie under the reference checkout, with the C
# driver oracle/ref_driver.cpp.  Test infrastructure only (a checker of the
# Green-tensor entries); nothing on the product path loads it.
#   usage: oracle/build_ref.sh [reference root]   (default /root/reference)
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${1:-/root/reference}"
SRC="$REF/piecewise_constant/src_vie/src_operators/mexcubatures"
if [ ! -d "$SRC" ]; then
  echo "build_ref.sh: no reference sources at $SRC" >&2
  exit 1
fi
mkdir -p "$HERE/_ref"
g++ -O2 -fPIC -shared -o "$HERE/_ref/libvpref.so.tmp" "$HERE/ref_driver.cpp" \
  "$SRC/cubature_nop_const.cpp" "$SRC/direct_ws_st_const.cpp" "$SRC/direct_ws_ea_const.cpp" \
  "$SRC/direct_ws_va_const.cpp"
mv "$HERE/_ref/libvpref.so.tmp" "$HERE/_ref/libvpref.so"
echo "$HERE/_ref/libvpref.so"
