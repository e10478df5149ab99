#!/bin/bash
# Test infrastructure only: compiles the reference's own, unmodified sources
# (read in place under the reference tree) together with oracle/ref_driver.cpp
# into oracle/_ref/libnf3d_ref.so.  The reference's CMake project cannot
# configure in this image (it needs pkg-config and TBB), so this script applies
# its per-file flags directly (CMakeLists.txt:57-59,63-69): -mavx for *_AVX.cpp,
# -mavx512f for *_AVX512.cpp, -DHAS_EXECUTION -DENABLE_PAR everywhere.
#
# FFT backend: the reference dlopens "libfftw3f_threads.so" (src/fftwlite.h:60).
# FFTW is not installed here; MKL's FFTW3 interface (libmkl_rt, in the image)
# exports the nine fftwf_* entry points the reference resolves, so
# oracle/_ref/libfftw3f_threads.so is a symlink to it, found through the
# driver library's $ORIGIN run path.
set -euo pipefail
REF=${NF3D_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
MKL=${NF3D_MKL:-/opt/conda/lib/libmkl_rt.so.1}
if [ ! -d "$REF/src" ]; then
  echo "reference tree not found at $REF; skipping oracle/_ref build" >&2
  exit 0
fi
mkdir -p "$OUT/obj"
CXX=${CXX:-g++}
FLAGS=(-O2 -std=gnu++17 -fPIC -DHAS_EXECUTION -DENABLE_PAR -w
       -I"$REF/src" -I"$REF/include/avisynth" -I"$REF/include/vapoursynth" -I"$REF/include/dualsynth")
objs=()
compile() {  # src extra-flags...
  local src=$1; shift
  local obj="$OUT/obj/$(basename "${src%.cpp}").o"
  "$CXX" "${FLAGS[@]}" "$@" -c "$src" -o "$obj"
  objs+=("$obj")
}
for s in buffer helper cpuid; do compile "$REF/src/$s.cpp"; done
for k in Apply Kalman Sharpen; do
  compile "$REF/src/code_impl/$k.cpp"
  compile "$REF/src/code_impl/${k}_SSE2.cpp"
  compile "$REF/src/code_impl/${k}_AVX.cpp" -mavx
  compile "$REF/src/code_impl/${k}_AVX512.cpp" -mavx512f
done
compile "$HERE/ref_driver.cpp"
"$CXX" -shared -o "$OUT/libnf3d_ref.so" "${objs[@]}" -Wl,--enable-new-dtags,-rpath,'$ORIGIN' -ldl -lpthread
ln -sf "$MKL" "$OUT/libfftw3f_threads.so"
# libmkl_rt loads its computational layers from its own directory as seen by
# the loader, i.e. from oracle/_ref through the symlink above.
for f in "$(dirname "$MKL")"/libmkl_*.so.1 "$(dirname "$MKL")"/libiomp5.so; do
  [ -e "$f" ] && ln -sf "$f" "$OUT/$(basename "$f")"
done
rm -rf "$OUT/obj"
echo "built $OUT/libnf3d_ref.so (FFT backend: $MKL)"
