#!/bin/sh
# Pulls the reference's plain-C host loops for the tile load balancer and the
# shared-queue offsets out of its .cu sources into include files under _ref/
# (never into the repository), so that ref_dump.cpp can compile them with g++.
#   src/main-cg.cu:596-769        balancer (non_each_block ... the packing loop)
#   src/main-cg-mixed.cu:2681-2709 shared-queue offsets
# Each range is checked against its first and last line so that a changed
# reference fails here instead of compiling something else.
# TEST INFRASTRUCTURE ONLY (fixture generation in the build container).
set -eu
REF=${1:?reference src dir}
OUT=${2:?output dir}
mkdir -p "$OUT"
check() {  # file line expected-substring
  got=$(sed -n "$2p" "$1")
  case "$got" in
    *"$3"*) ;;
    *) echo "ref_extract: $1:$2 is not '$3' (got '$got')" >&2; exit 1 ;;
  esac
}
check "$REF/main-cg.cu" 596 "int *non_each_block = (int *)malloc"
check "$REF/main-cg.cu" 770 "int *d_balance_tile_ptr_new;"
check "$REF/main-cg-mixed.cu" 2681 "int cnt_block = 0;"
check "$REF/main-cg-mixed.cu" 2709 "}"
check "$REF/main-cg-mixed.cu" 2711 "int *d_non_each_block_offset;"
sed -n '596,769p' "$REF/main-cg.cu" > "$OUT/ref_balance.inc"
sed -n '2681,2709p' "$REF/main-cg-mixed.cu" > "$OUT/ref_shared.inc"
