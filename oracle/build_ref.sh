#!/bin/bash
# Build the reference's own serial CPU kernels and its vendored gslib, from the
# sources where they lie under the reference tree, into oracle/_ref/ only.
# TEST INFRASTRUCTURE: the result (oracle/_ref/libnrsref.so) is the external
# check on the oracle's restatement and the source of tests/golden fixtures.
# Nothing is written into the reference tree; no reference source is copied.
#
# usage: oracle/build_ref.sh [REFERENCE_ROOT]   (default /root/reference)
set -euo pipefail
REF=${1:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/kernels/elliptic" ]; then
  echo "build_ref: reference tree not found at $REF; skipping" >&2
  exit 0
fi
mkdir -p "$OUT/obj"
K="$REF/kernels/elliptic"
GS="$REF/3rd_party/gslib/src"
CXX=${CXX:-g++}
CC=${CC:-gcc}
CXXFLAGS="-O2 -fPIC -fopenmp -D__NEKRS__OMP__ -x c++ -Ddlong=int -Dp_knl=0 -Dp_Nggeo=7
  -Dp_G00ID=0 -Dp_G01ID=1 -Dp_G11ID=2 -Dp_G12ID=3 -Dp_G02ID=4 -Dp_G22ID=5 -Dp_GWJID=6"
objs=()

# ellipticPartialAxCoeffHex3D.c (kernels/elliptic/ellipticPartialAxCoeffHex3D.c:2-97)
#   poisson, constant lambda (p_lambda=0)      : ref_ax_poisson_<T>_nq<Nq>
#   helmholtz, coefficient field (p_lambda=1)  : ref_ax_helm_<T>_nq<Nq>
for T in double float; do
  for nq in 2 3 4 5 6 8 10; do
    np=$((nq * nq * nq))
    o="$OUT/obj/ax_poisson_${T}_${nq}.o"
    $CXX $CXXFLAGS -Ddfloat=$T -Dp_Nq=$nq -Dp_Np=$np -Dp_lambda=0 -Dp_poisson=1 \
      "-DFUNC(a)=ref_ax_poisson_${T}_nq${nq}" -c "$K/ellipticPartialAxCoeffHex3D.c" -o "$o"
    objs+=("$o")
    o="$OUT/obj/ax_helm_${T}_${nq}.o"
    $CXX $CXXFLAGS -Ddfloat=$T -Dp_Nq=$nq -Dp_Np=$np -Dp_lambda=1 \
      "-DFUNC(a)=ref_ax_helm_${T}_nq${nq}" -c "$K/ellipticPartialAxCoeffHex3D.c" -o "$o"
    objs+=("$o")
  done
done

# ellipticBlockPartialAxCoeffHex3D.c (3 fields, coefficient fields)
for nq in 4 8; do
  np=$((nq * nq * nq))
  o="$OUT/obj/blockax_${nq}.o"
  $CXX $CXXFLAGS -Ddfloat=double -Dp_Nq=$nq -Dp_Np=$np -Dp_lambda=1 -Dp_Nfields=3 \
    "-DFUNC(a)=ref_blockax_nq${nq}" -c "$K/ellipticBlockPartialAxCoeffHex3D.c" -o "$o"
  objs+=("$o")
done

# ellipticBlockUpdatePCG.c (1 field)
o="$OUT/obj/updatepcg.o"
$CXX $CXXFLAGS -Ddfloat=double -Dp_Nfields=1 "-DFUNC(a)=ref_updatepcg" \
  -c "$K/ellipticBlockUpdatePCG.c" -o "$o"
objs+=("$o")

# ellipticPreconCoarsenHex3D.c / ellipticPreconProlongateHex3D.c (pfloat = double)
for pair in 8:4 8:2 4:2 8:6 6:4 10:4 10:6 6:2; do
  nf=${pair%%:*}; nc=${pair##*:}
  for kern in Coarsen Prolongate; do
    lk=$(echo $kern | tr 'A-Z' 'a-z')
    o="$OUT/obj/${lk}_${nf}_${nc}.o"
    $CXX $CXXFLAGS -Ddfloat=double -Dpfloat=double -Dp_NqFine=$nf -Dp_NqCoarse=$nc \
      -Dp_NpFine=$((nf * nf * nf)) -Dp_NpCoarse=$((nc * nc * nc)) \
      "-DFUNC(a)=ref_${lk}_${nf}_${nc}" -c "$K/ellipticPrecon${kern}Hex3D.c" -o "$o"
    objs+=("$o")
  done
done

# preFDM.c / fusedFDM.c (ASM: p_restrict=0, RAS: p_restrict=1) / postFDM.c, pfloat = float
#   ref_prefdm_nq<Nq>, ref_fusedfdm_{asm,ras}_nq<Nq>, ref_postfdm_nq<Nq>
for nq in 2 4 8; do
  nqe=$((nq + 2))
  F="-Dpfloat=float -Ddfloat=double -Dp_Nq=$nq -Dp_Nq_e=$nqe"
  o="$OUT/obj/prefdm_${nq}.o"
  $CXX $CXXFLAGS $F "-DFUNC(a)=ref_prefdm_nq${nq}" -c "$K/preFDM.c" -o "$o"
  objs+=("$o")
  for r in 0 1; do
    name=$([ $r = 1 ] && echo ras || echo asm)
    o="$OUT/obj/fusedfdm_${name}_${nq}.o"
    $CXX $CXXFLAGS $F -Dp_restrict=$r "-DFUNC(a)=ref_fusedfdm_${name}_nq${nq}" \
      -c "$K/fusedFDM.c" -o "$o"
    objs+=("$o")
  done
  o="$OUT/obj/postfdm_${nq}.o"
  $CXX $CXXFLAGS $F "-DFUNC(a)=ref_postfdm_nq${nq}" -c "$K/postFDM.c" -o "$o"
  objs+=("$o")
done

# gslib (serial: MPI undefined), 3rd_party/gslib/src
GSFLAGS="-O2 -fPIC -DUNDERSCORE -DGLOBAL_LONG_LONG -DUSE_NAIVE_BLAS -I$GS"
for f in gs sort sarray_transfer sarray_sort gs_local fail crystal comm tensor; do
  o="$OUT/obj/gslib_$f.o"
  $CC $GSFLAGS -c "$GS/$f.c" -o "$o"
  objs+=("$o")
done
o="$OUT/obj/ref_gs_wrap.o"
$CC $GSFLAGS -c "$HERE/ref_gs_wrap.c" -o "$o"
objs+=("$o")

$CXX -shared -fopenmp -o "$OUT/libnrsref.so" "${objs[@]}" -lm
echo "build_ref: wrote $OUT/libnrsref.so"
