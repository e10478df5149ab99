// Structures shared by the host engine (nf3d_engine.cpp) and the gfx950
// kernels (nf3d_kernels.hip).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace nf3d {

enum Mode : int {
  MODE_FWD = 0,      // gather + window + 2D r2c -> spectrum store (ring fill)
  MODE_W2D = 1,      // gather + r2c + Wiener2D (+sharpen/dehalo) + c2r
  MODE_W3D = 2,      // bt spectra (ring) + Wiener3D + sharpen + c2r
  MODE_KALMAN = 3,   // gather + r2c + Kalman (state in HBM) + sharpen + c2r
  MODE_SHARPEN = 4,  // gather + r2c + sharpen/dehalo + c2r
};

// Everything one launch needs about one plane of one frame.
struct BlockJob {
  // block geometry (src/fft3d_engine.h:112-136)
  int nox, noy, nb;
  int ow, oh;
  int mirw, mirh;
  int Wc, Hc;         // copied region (cropped plane size)
  int interlaced;
  // source pixels
  const uint8_t* src; // device plane base (row 0 of the uncropped plane)
  long long src_pitch;  // bytes
  int lc, tc;         // crop offsets in pixels
  int pix;            // bytes per sample: 1, 2 or 4 (float)
  float base;         // chroma centring subtracted in the gather (0 for luma and float)
  // tables (device)
  const float* wanx;  // [bw] analysis window along x (1 in the middle)
  const float* wany;  // [bh]
  const float2* grid; // [bh*outwidth] gridsample
  const float* wsharpen;
  const float* wdehalo;
  const float* pattern2d;
  const float* pattern3d;
  // spectral parameters (src/code_impl/code_impl.h:16-39)
  float sigma2;       // btcur * sigma^2 / norm
  float sigma2d;      // sigma^2 / norm (Kalman)
  float lowlimit;     // (beta-1)/beta
  float degrid;
  float sharpen, smin_n, smax_n, dehalo, ht2n, kratio2;
  int use_pattern, use_degrid;
  // spectra, layout [block][bh][outwidth] float2 (compact)
  float2* spec_out;          // MODE_FWD destination
  const float2* spec_in[5];  // MODE_W3D temporal inputs, oldest first
  const float2* spec_cur;    // the current frame's spectrum (degrid DC)
  float2* kal_last;
  float2* kal_covar;
  float2* kal_cproc;
  // synthesised blocks, layout [block][bh][bw] float
  float* blocks;
};

// Synthesis overlap-add of one plane (src/buffer.cpp:381-560 + 82-116).
struct OlaJob {
  int nox, noy, bw, bh, ow, oh, mirw, mirh, Wc, Hc, interlaced;
  const float* blocks;
  const float* wsynxl; const float* wsynxr; const float* wsynyl; const float* wsynyr;
  float norm;
  uint8_t* dst; long long dst_pitch; int lc, tc;
  int pix;        // 1, 2, 4
  int base;       // chroma planeBase (integer formats)
  int maxval;     // 2^bits - 1 (integer formats)
};

// Pattern estimation (src/helper.cpp:46-132).
struct PatternJob {
  int nox, noy, bh, outwidth;
  const float2* spec;
  const float2* grid;
  const float* pwin;
  float degrid;
  int px, py;          // fixed block, or 0,0 = search
  float* pattern2d;    // out [bh*outwidth]
  float* sums;         // scratch [nb]
  int* choice;         // out [2]: px, py
};

// Launchers; return hipSuccess or the launch error.  `bw`/`bh` select the
// compiled block shape; unsupported shapes return hipErrorInvalidValue.
hipError_t launch_block(int mode, int bt, int bw, int bh, const BlockJob& job, hipStream_t s);
hipError_t launch_ola(const OlaJob& job, hipStream_t s);
hipError_t launch_pattern(const PatternJob& job, hipStream_t s);
bool block_shape_supported(int bw, int bh);

}  // namespace nf3d
