// gfx950 kernels of the overlapped-block 3D frequency-domain denoiser.
//
// k_block<BW,BH,...>: one workgroup takes BPW blocks of one plane.
//   1. gather: the block pixels are read straight from the source plane in
//      HBM with reflect-101 cover addressing (no materialised cover), the
//      analysis window is applied with the reference's per-region product
//      order (src/buffer.cpp:137-378), and the floats are staged in LDS as
//      row pairs z = row(2p) + i*row(2p+1);
//   2. row r2c: one lane per row pair, a register FFT of length BW, split
//      into the two real-row spectra (k = 0..BW/2), written to LDS
//      column-major;
//   3. column FFT: one lane per spectral column, a register FFT of length BH;
//   4. the spectral pass of the mode (Wiener 2D/3D, Kalman, sharpen/dehalo,
//      degrid, pattern) on the lane's column, in registers;
//   5. column inverse FFT, LDS, row c2r (Hermitian pairs packed into one
//      complex transform per row pair), synthesised block rows to HBM.
// k_ola: the synthesis-windowed overlap-add of up to four blocks per pixel,
//   in the reference's float evaluation order (src/buffer.cpp:381-560),
//   rounded, clamped and stored to the destination plane.
#include "nf3d_fft.h"
#include "nf3d_internal.h"

namespace nf3d {

// ----------------------------------------------------------------- helpers
__device__ __forceinline__ int reflect101(int v, int n) {
  v = v < 0 ? -v : v;
  v = v >= n ? 2 * n - 2 - v : v;
  return v;
}

// ----------------------------------------------------------------- overlap
template <int PIX>
__global__ void __launch_bounds__(256) k_ola(OlaJob J) {
#pragma clang fp contract(off)
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= J.Wc || y >= J.Hc) return;
  const int stepx = J.bw - J.ow, stepy = J.bh - J.oh;
  int ry = y;
  if (J.interlaced) ry = (y & 1) == 0 ? (y >> 1) : (J.Hc >> 1) + ((J.Hc - 1 - y) >> 1);
  const int cy = ry + J.mirh, r = cy / stepy, i = cy - r * stepy;
  const int cx = x + J.mirw, c = cx / stepx, j = cx - c * stepx;
  const bool yb = i < J.oh && r >= 1, xb = j < J.ow && c >= 1;
  auto B = [&](int bx, int by, int ii, int jj) {
    return J.blocks[((size_t)(by * J.nox + bx) * J.bh + ii) * J.bw + jj];
  };
  float v;
  const float rounder = PIX == 4 ? 0.0f : 0.5f;
  const float base = (float)J.base;
  if (yb && xb) {
    float wyr = J.wsynyr[i] * J.norm, wyl = J.wsynyl[i] * J.norm;
    float wxr = J.wsynxr[j], wxl = J.wsynxl[j];
    float top = B(c - 1, r - 1, i + stepy, j + stepx) * wxr + B(c, r - 1, i + stepy, j) * wxl;
    float bot = B(c - 1, r, i, j + stepx) * wxr + B(c, r, i, j) * wxl;
    v = top * wyr + bot * wyl + rounder + base;
  } else if (yb) {
    float wyr = J.wsynyr[i] * J.norm, wyl = J.wsynyl[i] * J.norm;
    v = B(c, r - 1, i + stepy, j) * wyr + B(c, r, i, j) * wyl + rounder + base;
  } else if (xb) {
    v = (B(c - 1, r, i, j + stepx) * J.wsynxr[j] + B(c, r, i, j) * J.wsynxl[j]) * J.norm + rounder + base;
  } else {
    v = B(c, r, i, j) * J.norm + rounder + base;
  }
  uint8_t* row = J.dst + (long long)(y + J.tc) * J.dst_pitch;
  const int xx = x + J.lc;
  if constexpr (PIX == 4) {
    float o = v < 0.0f ? 0.0f : v;
    o = o > 1.0f ? 1.0f : o;
    reinterpret_cast<float*>(row)[xx] = o;
  } else {
    int q = (int)v;  // truncation toward zero, then clamp (buffer.cpp:406)
    q = q < 0 ? 0 : q;
    q = q > J.maxval ? J.maxval : q;
    if constexpr (PIX == 1) row[xx] = (uint8_t)q;
    else reinterpret_cast<uint16_t*>(row)[xx] = (uint16_t)q;
  }
}

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

// ---------------------------------------------------------------- dispatch
#define NF3D_SHAPES(X) X(8, 8) X(16, 16) X(24, 24) X(32, 32) X(48, 48) X(64, 64) X(32, 16) X(16, 32)

#define X(W, H) hipError_t launch_shape_##W##x##H(int mode, int bt, const BlockJob& job, hipStream_t s);
NF3D_SHAPES(X)
#undef X

bool block_shape_supported(int bw, int bh) {
#define X(W, H) if (bw == W && bh == H) return true;
  NF3D_SHAPES(X)
#undef X
  return false;
}

hipError_t launch_block(int mode, int bt, int bw, int bh, const BlockJob& job, hipStream_t s) {
#define X(W, H) if (bw == W && bh == H) return launch_shape_##W##x##H(mode, bt, job, s);
  NF3D_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_ola(const OlaJob& J, hipStream_t s) {
  dim3 grid((J.Wc + 255) / 256, J.Hc);
  if (J.pix == 1) hipLaunchKernelGGL(k_ola<1>, grid, dim3(256), 0, s, J);
  else if (J.pix == 2) hipLaunchKernelGGL(k_ola<2>, grid, dim3(256), 0, s, J);
  else hipLaunchKernelGGL(k_ola<4>, grid, dim3(256), 0, s, J);
  return hipGetLastError();
}

hipError_t launch_pattern(const PatternJob& J, hipStream_t s) {
  int nb = J.nox * J.noy;
  hipLaunchKernelGGL(k_pattern_sums, dim3((nb + 127) / 128), dim3(128), 0, s, J);
  hipLaunchKernelGGL(k_pattern_set, dim3(1), dim3(256), 0, s, J);
  return hipGetLastError();
}

}  // namespace nf3d
