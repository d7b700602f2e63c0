#!/bin/bash
# Builds one reference checker executable (see Makefile in this folder).
# usage: build_one.sh <variant r8|shipped> <grid JxKxL> <reference tree> <output dir> <flang>
set -euo pipefail
variant=$1; grid=$2; ref=$3; out=$4; flang=${5:-flang}
IFS=x read -r J K L <<< "$grid"
here=$(cd "$(dirname "$0")" && pwd)
common="-fopenmp -cpp -mcmodel=medium"
case $variant in
  r8)      flags="-O3 $common -fdefault-real-8 -fdefault-double-8"; bdyflags="$flags" ;;
  shipped) flags="-O3 $common";                  bdyflags="-O0 $common" ;;
  *) echo "unknown variant $variant" >&2; exit 2 ;;
esac
dir=$out/${variant}_${grid}
mkdir -p "$dir/inc"
# configured copy of the reference's grid header: only the five grid PARAMETERs change
sed -E -e "s/\bjmax=256\b/jmax=$J/" -e "s/\bkmax=128\b/kmax=$K/" \
       -e "s/\blmax=256\b/lmax=$L/" -e "s/\bpot3jmax=256\b/pot3jmax=$J/" \
       -e "s/\bpot3kmax=128\b/pot3kmax=$K/" \
       "$ref/headers/hydroparam.h" > "$dir/inc/hydroparam.h"
inc="-I$dir/inc -I$ref/headers"
cd "$dir"
$flang $flags    -ffixed-form $inc -c "$ref/pot/pot3.f"     -o pot3.o
$flang $flags    -ffixed-form $inc -c "$ref/pot/fft.f"      -o fft.o
$flang $flags    -ffixed-form $inc -c "$ref/pot/blktri.f"   -o blktri.o
$flang $bdyflags -ffixed-form $inc -c "$ref/pot/boundary.f" -o boundary.o
$flang $flags    -ffree-form  $inc -c "$here/refdriver.F90" -o refdriver.o
# GNU ld: lld rejects the 2.9 GB of static COMMON at the largest grid
$flang $flags -fuse-ld=bfd -Wl,--no-relax -o refpot refdriver.o pot3.o fft.o blktri.o boundary.o
echo "built $dir/refpot"
