// One block shape of k_tile / k_seam: compiled once per shape with
// -DNF3D_BW=<w> -DNF3D_BH=<h> so the shapes build in parallel; the dispatcher
// (nf3d_kernels.hip) calls the launcher by name.
#include "nf3d_tile.h"

#if !defined(NF3D_BW) || !defined(NF3D_BH)
#error "compile with -DNF3D_BW=<w> -DNF3D_BH=<h>"
#endif

#define NF3D_NAME2(a, b) launch_tile_##a##x##b
#define NF3D_NAME(a, b) NF3D_NAME2(a, b)

namespace nf3d {
hipError_t NF3D_NAME(NF3D_BW, NF3D_BH)(int mode, int bt, int onchip, const TileLaunch& l, hipStream_t s) {
  return launch_tile_shape<NF3D_BW, NF3D_BH>(mode, bt, onchip, l, s);
}
}  // namespace nf3d
