#!/usr/bin/env bash
# Builds oracle/_ref/refdrv from the reference's own sources where they lie
# under /root/reference (nothing is copied) plus oracle/ref_driver.cpp.
#
# Flags follow the reference build (-O3 -DNDEBUG -fopenmp, no -march, so no FMA
# contraction).  The reference's partial_reference.cpp needs a header that its
# CMake configure step generates; it is not compiled, and --gc-sections drops the
# cache-key code paths that would reference it.  MPI (MPICH in /opt/conda) and
# OpenSSL are the image's own libraries; libstdc++ is linked statically so the
# binary runs without conda's older libstdc++ on the library path.
set -euo pipefail
REF=${REFERENCE_ROOT:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/src" ]; then
  echo "build_ref: $REF not present; skipping reference build" >&2
  exit 0
fi
MPI_INC=${MPI_INC:-/opt/conda/include}
MPI_LIB=${MPI_LIB:-/opt/conda/lib}
mkdir -p "$OUT/obj"
S="$REF/src"
FLAGS="-std=c++17 -O3 -DNDEBUG -fopenmp -ffunction-sections -fdata-sections -DSPGEMM_MPI_CHUNK_LIMIT=2147483647 \
 -I$S/common -I$S/baselines -I$S/trident/common -I$MPI_INC -w"
SRCS="$S/common/spgemm_common.cpp $S/common/matrix_market.cpp $S/common/full_reference.cpp \
 $S/baselines/spgemm_exchange.cpp $S/trident/common/matrix_blocks.cpp $S/trident/common/process_grid.cpp \
 $HERE/ref_driver.cpp"
objs=()
for f in $SRCS; do
  o="$OUT/obj/$(basename "$f" .cpp).o"
  g++ $FLAGS -c "$f" -o "$o"
  objs+=("$o")
done
g++ -fopenmp -Wl,--gc-sections "${objs[@]}" -L"$MPI_LIB" -Wl,-rpath,"$MPI_LIB" -lmpi -lcrypto \
  -static-libstdc++ -o "$OUT/refdrv"
echo "build_ref: built $OUT/refdrv"
