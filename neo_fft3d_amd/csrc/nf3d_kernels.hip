// gfx950 kernels of the denoiser other than the per-shape tile kernels
// (nf3d_tile.h, compiled per block shape from nf3d_shape.hip): the pattern
// estimator and the shape dispatcher.
#include "nf3d_fft.h"
#include "nf3d_internal.h"

namespace nf3d {

// ----------------------------------------------------------------- pattern
// Sum of the windowed, degrid-corrected power of each candidate block, in
// the reference's sequential float order (src/helper.cpp:57-89).
__global__ void k_pattern_sums(PatternJob J) {
#pragma clang fp contract(off)
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  int nb = J.nox * J.noy;
  if (b >= nb) return;
  int by = b / J.nox, bx = b - by * J.nox;
  if (by < 2 || by >= J.noy - 2 || bx < 2 || bx >= J.nox - 2) { J.sums[b] = 3.0e38f; return; }
  const float2* blk = J.spec + (size_t)b * J.bh * J.outwidth;
  float gcur = J.degrid * blk[0].x / J.grid[0].x;
  float s = 0.0f;
  for (int k = 0; k < J.bh * J.outwidth; k++) {
    float2 g = J.grid[k];
    float c0 = blk[k].x - gcur * g.x, c1 = blk[k].y - gcur * g.y;
    float psd = c0 * c0 + c1 * c1;
    s += psd * J.pwin[k];
  }
  J.sums[b] = s;
}

// First strict minimum in row-major block order, then the pattern of the
// chosen block (src/helper.cpp:82-88, 111-130).  One workgroup.
__global__ void __launch_bounds__(256) k_pattern_set(PatternJob J) {
#pragma clang fp contract(off)
  __shared__ int choice[2];
  if (threadIdx.x == 0) {
    int px = J.px, py = J.py;
    if (px == 0 && py == 0) {
      float best = 1e15f;
      for (int by = 2; by < J.noy - 2; by++)
        for (int bx = 2; bx < J.nox - 2; bx++) {
          float s = J.sums[by * J.nox + bx];
          if (s < best) { best = s; px = bx; py = by; }
        }
    }
    choice[0] = px;
    choice[1] = py;
    J.choice[0] = px;
    J.choice[1] = py;
  }
  __syncthreads();
  const float2* blk = J.spec + ((size_t)choice[1] * J.nox + choice[0]) * J.bh * J.outwidth;
  float gcur = J.degrid * blk[0].x / J.grid[0].x;
  for (int k = threadIdx.x; k < J.bh * J.outwidth; k += blockDim.x) {
    float2 g = J.grid[k];
    float c0 = blk[k].x - gcur * g.x, c1 = blk[k].y - gcur * g.y;
    float psd = c0 * c0 + c1 * c1;
    J.pattern2d[k] = psd * J.pwin[k];
  }
}

// Spatially varying noise pattern (no reference counterpart; it generalises
// FindPatternBlock / SetPattern, src/helper.cpp:46-132, to regions): the
// block grid is cut into regions of rx x ry blocks, and each region takes the
// pattern of its own first strict minimum of the same sums over the blocks
// of the reference's search range (blocks 2..nox-3, 2..noy-3) that lie in
// it -- or, for a region with none of them, the block chosen for the whole
// plane (J.choice, written by k_pattern_set).  The pattern goes to every
// block of the region: out2d = |X - g G|^2 pwin, out3d = out2d * mult3d.
// One workgroup per region.
__global__ void __launch_bounds__(256) k_pattern_regions(PatternJob J, int rx, int ry, float* out2d, float* out3d,
                                                         float mult3d) {
#pragma clang fp contract(off)
  __shared__ int choice[2];
  const int nqx = (J.nox + rx - 1) / rx;
  const int qy = blockIdx.x / nqx, qx = blockIdx.x - qy * nqx;
  const int x0 = qx * rx, x1 = min(J.nox, x0 + rx), y0 = qy * ry, y1 = min(J.noy, y0 + ry);
  if (threadIdx.x == 0) {
    int px = -1, py = -1;
    float best = 1e15f;
    for (int by = max(y0, 2); by < min(y1, J.noy - 2); by++)
      for (int bx = max(x0, 2); bx < min(x1, J.nox - 2); bx++) {
        const float v = J.sums[by * J.nox + bx];
        if (v < best) { best = v; px = bx; py = by; }
      }
    if (px < 0) {
      px = J.choice[0];
      py = J.choice[1];
    }
    choice[0] = px;
    choice[1] = py;
  }
  __syncthreads();
  const int n = J.bh * J.outwidth;
  const float2* blk = J.spec + ((size_t)choice[1] * J.nox + choice[0]) * n;
  const float gcur = J.degrid * blk[0].x / J.grid[0].x;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const float2 g = J.grid[k];
    const float c0 = blk[k].x - gcur * g.x, c1 = blk[k].y - gcur * g.y;
    const float v = (c0 * c0 + c1 * c1) * J.pwin[k];
    const float v3 = v * mult3d;
    for (int by = y0; by < y1; by++)
      for (int bx = x0; bx < x1; bx++) {
        const size_t o = ((size_t)by * J.nox + bx) * n + k;
        out2d[o] = v;
        out3d[o] = v3;
      }
  }
}

// ---------------------------------------------------------------- dispatch
// Block shapes compiled for k_tile (one translation unit each).
#define NF3D_SHAPES(X) X(8, 8) X(16, 16) X(24, 24) X(32, 32) X(48, 48) X(64, 64) X(32, 16) X(16, 32)

#define X(W, H)                                                                                          \
  hipError_t launch_tile_##W##x##H(int mode, int bt, int onchip, int mf, const TileLaunch& l, hipStream_t s); \
  hipError_t launch_run_##W##x##H(int mode, int bt, int shp, int mf, const RunBatch& l, hipStream_t s); \
  hipError_t launch_seam_run_##W##x##H(const RunBatch& l, hipStream_t s);
NF3D_SHAPES(X)
#undef X

bool block_shape_supported(int bw, int bh) {
#define X(W, H) if (bw == W && bh == H) return true;
  NF3D_SHAPES(X)
#undef X
  return false;
}

int tile_blocks_per_wg(int bw, int bh) {
  // TShape<BW,BH,128>::BPW: 128 lanes / max(bw/2+1, bh/2)
  int cols = bw / 2 + 1, pairs = bh / 2;
  return 128 / (cols > pairs ? cols : pairs);
}

bool mfma_dft_supported(int bw, int bh) { return bw == bh && (bw == 32 || bw == 48 || bw == 64); }

hipError_t launch_tile(int mode, int bt, int onchip, int mf, int bw, int bh, const TileLaunch& l, hipStream_t s) {
  if (mf && !mfma_dft_supported(bw, bh)) return hipErrorInvalidValue;
#define X(W, H) if (bw == W && bh == H) return launch_tile_##W##x##H(mode, bt, onchip, mf, l, s);
  NF3D_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_run(int mode, int bt, int shp, int mf, int bw, int bh, const RunBatch& l, hipStream_t s) {
  if (mf && !mfma_dft_supported(bw, bh)) return hipErrorInvalidValue;
#define X(W, H) if (bw == W && bh == H) return launch_run_##W##x##H(mode, bt, shp, mf, l, s);
  NF3D_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_seam_run(int bw, int bh, const RunBatch& l, hipStream_t s) {
#define X(W, H) if (bw == W && bh == H) return launch_seam_run_##W##x##H(l, s);
  NF3D_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_pattern(const PatternJob& J, hipStream_t s) {
  int nb = J.nox * J.noy;
  hipLaunchKernelGGL(k_pattern_sums, dim3((nb + 127) / 128), dim3(128), 0, s, J);
  hipLaunchKernelGGL(k_pattern_set, dim3(1), dim3(256), 0, s, J);
  return hipGetLastError();
}

hipError_t launch_pattern_regions(const PatternJob& J, int rx, int ry, float* out2d, float* out3d, float mult3d,
                                  hipStream_t s) {
  if (rx < 1 || ry < 1) return hipErrorInvalidValue;
  const int nq = ((J.nox + rx - 1) / rx) * ((J.noy + ry - 1) / ry);
  hipLaunchKernelGGL(k_pattern_regions, dim3(nq), dim3(256), 0, s, J, rx, ry, out2d, out3d, mult3d);
  return hipGetLastError();
}

}  // namespace nf3d
