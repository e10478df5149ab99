// One block shape of the kernels, compiled per shape with -DNF3D_BW=<w>
// -DNF3D_BH=<h> and once per part: -DNF3D_PART_TILE (k_tile / k_seam, one
// frame per launch) or -DNF3D_PART_RUN (k_run / k_seam_run, frame loop), so
// the units build in parallel; the dispatcher (nf3d_kernels.hip) calls the
// launchers by name.
#include "nf3d_run.h"

#if !defined(NF3D_BW) || !defined(NF3D_BH)
#error "compile with -DNF3D_BW=<w> -DNF3D_BH=<h>"
#endif

#define NF3D_NAME2(a, b) launch_tile_##a##x##b
#define NF3D_NAME(a, b) NF3D_NAME2(a, b)
#define NF3D_RNAME2(a, b) launch_run_##a##x##b
#define NF3D_RNAME(a, b) NF3D_RNAME2(a, b)
#define NF3D_SNAME2(a, b) launch_seam_run_##a##x##b
#define NF3D_SNAME(a, b) NF3D_SNAME2(a, b)

namespace nf3d {
#if defined(NF3D_PART_TILE)
hipError_t NF3D_NAME(NF3D_BW, NF3D_BH)(int mode, int bt, int onchip, int mf, const TileLaunch& l, hipStream_t s) {
  return launch_tile_shape<NF3D_BW, NF3D_BH>(mode, bt, onchip, mf, l, s);
}
#elif defined(NF3D_PART_RUN)
hipError_t NF3D_RNAME(NF3D_BW, NF3D_BH)(int mode, int bt, int shp, int mf, const RunBatch& l, hipStream_t s) {
  return launch_run_shape<NF3D_BW, NF3D_BH>(mode, bt, shp, mf, l, s);
}
hipError_t NF3D_SNAME(NF3D_BW, NF3D_BH)(const RunBatch& l, hipStream_t s) {
  return launch_seam_run_shape<NF3D_BW, NF3D_BH>(l, s);
}
#else
#error "compile with -DNF3D_PART_TILE or -DNF3D_PART_RUN"
#endif
}  // namespace nf3d
