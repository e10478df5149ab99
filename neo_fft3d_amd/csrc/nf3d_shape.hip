// One block shape of k_block: compiled once per shape with -DNF3D_BW=<w>
// -DNF3D_BH=<h> so the shapes build in parallel; the dispatcher
// (nf3d_kernels.hip) calls the launcher by name.
#include "nf3d_block.h"

#if !defined(NF3D_BW) || !defined(NF3D_BH)
#error "compile with -DNF3D_BW=<w> -DNF3D_BH=<h>"
#endif

#define NF3D_CAT2(a, b, c) launch_shape_##a##x##b##c
#define NF3D_CAT(a, b) NF3D_CAT2(a, b, )

namespace nf3d {
hipError_t NF3D_CAT(NF3D_BW, NF3D_BH)(int mode, int bt, const BlockJob& job, hipStream_t s) {
  return launch_shape<NF3D_BW, NF3D_BH>(mode, bt, job, s);
}
}  // namespace nf3d
