// Structures shared by the host engine (nf3d_engine.cpp) and the gfx950
// kernels (nf3d_kernels.hip).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace nf3d {

enum Mode : int {
  MODE_FWD = 0,      // gather + window + 2D r2c -> spectrum store (ring fill)
  MODE_W2D = 1,      // gather + r2c + Wiener2D (+sharpen/dehalo) + c2r
  MODE_W3D = 2,      // bt spectra (ring, newest optionally on chip) + Wiener3D + sharpen + c2r
  MODE_KALMAN = 3,   // gather + r2c + Kalman (state in HBM) + sharpen + c2r
  MODE_SHARPEN = 4,  // gather + r2c + sharpen/dehalo + c2r
};

// Everything one launch needs about one plane of one frame.
struct PlaneJob {
  // block geometry (src/fft3d_engine.h:112-136)
  int nox, noy, nb;
  int ow, oh;
  int mirw, mirh;
  int Wc, Hc;         // copied region (cropped plane size)
  int interlaced;
  int ntx;            // tiles per block row: ceil(nox / blocks per workgroup)
  int wg_begin;       // first workgroup of this plane in the launch
  int seam_row_begin; // first k_seam grid row of this plane
  // source pixels
  const uint8_t* src; // device plane base (row 0 of the uncropped plane)
  long long src_pitch;  // bytes
  int lc, tc;         // crop offsets in pixels
  int pix;            // bytes per sample: 1, 2 or 4 (float)
  float base;         // chroma centring subtracted in the gather (0 for luma and float)
  // destination
  uint8_t* dst;
  long long dst_pitch;
  int obase;          // planeBase added back on output (integer chroma)
  int maxval;         // 2^bits - 1 (integer formats)
  float norm;         // 1 / (bw*bh)
  float* seam;        // [4][Hc][Wc] partial sums of bands shared by tiles
  long long seam_plane;  // Hc * Wc
  // tables (device)
  const float* wanx;  // [bw] analysis window along x (1 in the middle)
  const float* wany;  // [bh]
  const float* wsynxl; const float* wsynxr; const float* wsynyl; const float* wsynyr;
  const float2* grid; // [bh*outwidth] gridsample
  const float* wsharpen;
  const float* wdehalo;
  const float* pattern2d;
  const float* pattern3d;
  const int* keep_choice;  // pattern display: keep only block (keep_choice[0], keep_choice[1]); null = off
  // spectral parameters (src/code_impl/code_impl.h:16-39)
  float sigma2;       // btcur * sigma^2 / norm
  float sigma2d;      // sigma^2 / norm (Kalman)
  float lowlimit;     // (beta-1)/beta
  float degrid;
  float sharpen, smin_n, smax_n, dehalo, ht2n, kratio2;
  int use_pattern, use_degrid;
  int pattern_per_block;  // pattern2d / pattern3d hold one [bh][outwidth] pattern per block (learned PSD)
  // matrix-core DFT (32x32 blocks, per-frame kernel): constant tiles and the
  // power-of-two operand scales of the forward and inverse transforms
  const void* dft_tab;
  float dft_sx, dft_sz;
  // spectra, layout [block][bh][outwidth] float2 (compact)
  float2* spec_out;          // where this launch's forward spectra are stored (or null)
  const float2* spec_in[5];  // Wiener-3D window, oldest first (on-chip newest: unused)
  float2* kal_last;
  float2* kal_covar;
  float2* kal_cproc;
};

// One launch: up to three planes of one frame (same block shape).
struct TileLaunch {
  PlaneJob job[3];
  int njobs;
  int total_wg;
  int total_seam_rows;
  int max_wc;
};

// Frame-loop launch (k_run): one plane, up to RUN_MAX_FRAMES consecutive
// frames, every tile of the plane.
constexpr int RUN_MAX_FRAMES = 16;

// Specialisation of the sharpen / dehalo factors in k_run's spectral pass.
enum Shp : int {
  SHP_NONE = 0,     // sharpen == 0 and dehalo == 0
  SHP_SHARPEN = 1,  // dehalo == 0
  SHP_GENERAL = 3,  // dehalo != 0 (sharpen factor times dehalo factor; 1 exactly when sharpen == 0)
};

// One job of a frame-loop launch: a run of frames of one plane of one clip.
struct RunLaunch {
  PlaneJob J;                           // geometry, tables, parameters; J.seam = per-frame seam buffers
  const uint8_t* src[RUN_MAX_FRAMES];   // plane transformed on chip at step i (frame m0+i, +(bt-1)/2 for Wiener 3D)
  uint8_t* dst[RUN_MAX_FRAMES];         // output plane of frame m0+i
  float2* ring;                         // Wiener 3D: bt slots of spec_elems, frame k in slot k % bt
  long long spec_elems;
  int m0, nframes;
  int fpw;                              // consecutive frames filtered by one workgroup
  int ntiles;                           // tiles of the plane (noy * ntx)
  int wg_begin;                         // first workgroup of this job in the launch
};

// Jobs of one launch (same block shape, mode, bt and sharpen case): the
// planes of a clip, or of several clips.
constexpr int RUN_MAX_JOBS = 6;
struct RunBatch {
  RunLaunch job[RUN_MAX_JOBS];
  int njobs;
  int total_wg;    // sum of the jobs' workgroups
  int max_hc;      // largest plane height (seam grid)
  int max_frames;  // largest nframes (seam grid)
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
// launch_tile runs k_tile and, for the filtering modes, k_seam.
hipError_t launch_tile(int mode, int bt, int onchip, int mf, int bw, int bh, const TileLaunch& l, hipStream_t s);
// Block shapes with a matrix-core DFT variant of k_tile (mf = 1).
bool mfma_dft_supported(int bw, int bh);
hipError_t launch_pattern(const PatternJob& job, hipStream_t s);
// Per-region patterns after launch_pattern (which fills sums and choice):
// every block gets its rx x ry region's pattern, 2D and times mult3d.
hipError_t launch_pattern_regions(const PatternJob& job, int rx, int ry, float* out2d, float* out3d, float mult3d,
                                  hipStream_t s);
// DFT-as-GEMM precision form (nf3d_mfma_dft.hip): n x n blocks, n = 32 or 64.
hipError_t launch_dft_mfma(const void* in, void* out, int nb, int n, int inverse, float scale, hipStream_t s);
// k_run over the batch's jobs; `shp` is a Shp value.  launch_seam_run
// finishes the seams of their frames (after the k_run, same stream).
hipError_t launch_run(int mode, int bt, int shp, int mf, int bw, int bh, const RunBatch& b, hipStream_t s);
hipError_t launch_seam_run(int bw, int bh, const RunBatch& b, hipStream_t s);
bool block_shape_supported(int bw, int bh);
int tile_blocks_per_wg(int bw, int bh);

}  // namespace nf3d
