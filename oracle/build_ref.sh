#!/bin/bash
# Builds the PDAC reference solver (serial, -D__SCALAR) from its Fortran and C
# sources in /root/reference with flang, writing objects, modules and the
# executable only under oracle/_ref/.  The reference's own Makefile is not used:
# the module dependency order is resolved here by repeated passes.
# The resulting pdac.x is a test instrument: it is run in the build container
# to write golden fixtures and is never linked into the product.
set -u
REF=${PDAC_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
FC=${FC:-flang}
CC=${CC:-gcc}
[ -d "$REF" ] || { echo "reference sources not found at $REF"; exit 0; }
mkdir -p "$OUT/obj" "$OUT/mod"
FFLAGS="-O2 -ffree-form -w"
# solver sources of pdac.x (the post-processor sources are not needed)
SRCS="arrfil atmosphere blunt bdry control decomp dens derived dimensions dome drag
environment eosg eosl explicit_solver ftem flux_u flux_v flux_w flux_sc flux_m gas ghost grid
hcapgs hotc hvs htilde kinds immb indijk input interp inoutflow io_parallel io_restart iter kb07ad
limiters letter matrix mean_fields mixture outp parallel particles press prog reactions rannum
residuals roughness samplert setup sink subscr temp tilde time topo turbo types velocity vent
visc ygas iotk_module io_files pdac"
COMM_F="barrier startup hangup bcast gather scatter sendrecv reduce error wtime
myflush isend recv wait irecv"
COMM_FF="commutil localdim"
COMM_C="cptimer memstat"
cd "$OUT/obj"
for f in $COMM_F; do
  [ -f $f.o ] || $FC -c -cpp -D__SCALAR -O2 -w -J "$OUT/mod" -o $f.o "$REF/comm/$f.F" || exit 1
done
for f in $COMM_FF; do
  [ -f $f.o ] || $FC -c -O2 -w -J "$OUT/mod" -o $f.o "$REF/comm/$f.f" || exit 1
done
for f in $COMM_C; do
  [ -f $f.o ] || $CC -c -O2 -D__SCALAR -D__BEOWULF -o $f.o "$REF/comm/$f.c" || exit 1
done
todo="$SRCS"
pass=0
while [ -n "$(echo $todo)" ]; do
  pass=$((pass+1)); left=""; progress=0
  for f in $todo; do
    if [ -f $f.o ]; then continue; fi
    if $FC -c $FFLAGS -J "$OUT/mod" -o $f.o "$REF/$f.f" 2>/dev/null; then
      progress=1
    else
      left="$left $f"
    fi
  done
  todo="$left"
  if [ $progress -eq 0 ] && [ -n "$(echo $todo)" ]; then
    echo "cannot resolve: $todo"
    for f in $todo; do $FC -c $FFLAGS -J "$OUT/mod" -o $f.o "$REF/$f.f"; done
    exit 1
  fi
done
$FC -o "$OUT/pdac.x" $(for f in $SRCS; do echo $f.o; done) \
   $(for f in $COMM_F $COMM_FF $COMM_C; do echo $f.o; done) || exit 1
echo "built $OUT/pdac.x in $pass passes"
