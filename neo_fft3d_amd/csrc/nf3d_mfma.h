// 32x32 block DFTs as dense DFT-matrix products on the gfx950 matrix cores,
// shared by the standalone transforms (nf3d_mfma_dft.hip) and the filter's
// kernels (k_run / k_tile with the matrix-core DFT).
//
// Every real operand is split into two FP16 halves, x = hi + lo with
// hi = fp16(x) and lo = fp16(x - hi), and a product is taken as
// hi*hi + hi*lo + lo*hi with FP32 accumulation (v_mfma_f32_16x16x32_f16,
// three per 16x16 tile, the whole K = 32 in one instruction): about 22
// significant bits per operand against FP32's 24, so the transforms land
// within a few FP32 ulps of an FFT.  Operands are scaled by powers of two
// (exact) to stay inside the FP16 range; the output is scaled back exactly.
//
// One wavefront owns one block.  Layouts of the 16x16x32 MFMA (lane l,
// g = l / 16; checked on the GPU by a probe):
//   A (16x32, row-major M x K): lane l holds A[l % 16][K(g, i)], i < 8
//   B (K x N):                  lane l holds B[K(g, i)][l % 16]
//   D (M x N):                  lane l holds D[4g + i][l % 16], i < 4
// A product's K order is free as long as A and B agree.  Operands loaded
// from memory use K(g, i) = 8g + i.  A product whose B operand is the pair
// of D tiles (rows 0..15 and 16..31) of the previous product uses
// K2(g, i) = 4g + i for i < 4 and 16 + 4g + (i - 4) for i >= 4: the lane's
// four values of each D tile, side by side, are then directly its B
// operand, and the constant A tiles of that product are stored in K2 order.
// The chains:
//   forward  Y = X C        (C: 32 x 64, columns 0..16 cos, 32..48 -sin)
//            Z = F Y        (F = Fr + i Fi, Fr = cos, Fi = -sin; 32 x 32)
//            as P = Fr [Yr|Yi], Q = Fi [Yr|Yi]: Zr = P_r - Q_i, Zi = P_i + Q_r
//   inverse  Y^T = Z^T conj(F)            (per spectral column, all rows)
//            x^T = Dr^T Y^T_r + Dn^T Y^T_i (Dr = w cos, Dn = -w sin, w = 1 at
//            k = 0 and 16, else 2: the c2r of the last dimension, the
//            imaginary parts of columns 0 and 16 dropped as FFTW's c2r does)
// Each direction is split in two: the operand loads (into registers) and the
// products (with the results stored), so a caller whose source and
// destination share memory can load every block first.
#pragma once
#include <hip/hip_runtime.h>

namespace nf3d {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

struct H2 {  // four values split (a D tile's share of a lane)
  half4 hi, lo;
};
struct H8 {  // eight values split (an A or B operand with K = 32)
  half8 hi, lo;
};
typedef H8 DftTile;

__device__ __forceinline__ H2 split4(f4 v, float s) {
  H2 r;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const float x = v[i] * s;
    const _Float16 h = (_Float16)x;
    r.hi[i] = h;
    r.lo[i] = (_Float16)(x - (float)h);
  }
  return r;
}

// Two D-tile shares (rows 0..15 and 16..31) as one K2-ordered operand.
__device__ __forceinline__ H8 cat(const H2& a, const H2& b) {
  H8 r;
  r.hi = __builtin_shufflevector(a.hi, b.hi, 0, 1, 2, 3, 4, 5, 6, 7);
  r.lo = __builtin_shufflevector(a.lo, b.lo, 0, 1, 2, 3, 4, 5, 6, 7);
  return r;
}

__device__ __forceinline__ f4 mma3(const H8& a, const H8& b, f4 c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.lo, b.hi, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.hi, b.lo, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.hi, b.hi, c, 0, 0, 0);
  return c;
}

// Constant tiles, [tile][lane] of H8.
struct DftTab {
  static constexpr int C = 0;    // forward B tiles of C (K order 8g+i): n -> n, 4 tiles
  static constexpr int FR = 4;   // forward A tiles of Fr (K2 order): m -> m, 2 tiles
  static constexpr int FI = 6;   // forward A tiles of Fi
  static constexpr int BR = 8;   // inverse B tiles of Fr (K order 8g+i): n -> n, 2 tiles
  static constexpr int BI = 10;  // inverse B tiles of Fi
  static constexpr int BN = 12;  // inverse B tiles of -Fi
  static constexpr int DR = 14;  // inverse A tiles of Dr^T (K2 order): m -> m, 2 tiles
  static constexpr int DN = 16;  // inverse A tiles of Dn^T
  static constexpr int COUNT = 18;
};

__device__ __forceinline__ H8 tile(const H8* T, int t, int lane) { return T[t * 64 + lane]; }

// ---------------------------------------------------------------- forward
struct DftIn {
  H8 a[2];  // A tiles of x: row tile m, all 32 columns
};

__device__ __forceinline__ H8 split8(const float (&v)[8], float s) {
  f4 v0 = {v[0], v[1], v[2], v[3]}, v1 = {v[4], v[5], v[6], v[7]};
  return cat(split4(v0, s), split4(v1, s));
}

// x(r, c) -> operand tiles, times sx.
template <class X>
__device__ __forceinline__ DftIn dft32_fwd_load(X x, float sx, int lane) {
  const int lr = lane & 15, lg = lane >> 4;
  DftIn d;
#pragma unroll
  for (int m = 0; m < 2; m++) {
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = x(16 * m + lr, 8 * lg + i);
    d.a[m] = split8(v, sx);
  }
  return d;
}

// Z = F X C, handed to store(r, k, z) for k < 17 (each (r, k) once).
template <class Z>
__device__ __forceinline__ void dft32_fwd_compute(const DftIn& d, float sx, const H8* T, int lane, Z store) {
  const int lr = lane & 15, lg = lane >> 4;
  H2 yb[2][4];
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const f4 y = mma3(d.a[m], tile(T, DftTab::C + n, lane), f4{0.0f, 0.0f, 0.0f, 0.0f});
      yb[m][n] = split4(y, 1.0f / 32.0f);  // |Y| <= 32 max|x|
    }
  const float back = 32.0f / sx;
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int n = 0; n < 2; n++) {
      const H8 fr = tile(T, DftTab::FR + m, lane), fi = tile(T, DftTab::FI + m, lane);
      const H8 yr = cat(yb[0][n], yb[1][n]), yi = cat(yb[0][n + 2], yb[1][n + 2]);
      const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
      const f4 pr = mma3(fr, yr, zero);  // Fr Yr
      const f4 pi = mma3(fr, yi, zero);  // Fr Yi
      const f4 qr = mma3(fi, yr, zero);  // Fi Yr
      const f4 qi = mma3(fi, yi, zero);  // Fi Yi
      const int k = 16 * n + lr;
      if (k < 17) {
#pragma unroll
        for (int i = 0; i < 4; i++)
          store(16 * m + 4 * lg + i, k, make_float2((pr[i] - qi[i]) * back, (pi[i] + qr[i]) * back));
      }
    }
}

// ---------------------------------------------------------------- inverse
struct DftSpec {
  H8 ar[2], ai[2];  // A tiles of Z^T (real, imaginary): spectral-column tile m, all 32 rows
};

// z(r, k) (k < 17) -> operand tiles, times sz.
template <class Z>
__device__ __forceinline__ DftSpec dft32_inv_load(Z z, float sz, int lane) {
  const int lr = lane & 15, lg = lane >> 4;
  DftSpec d;
#pragma unroll
  for (int m = 0; m < 2; m++) {
    const int k = 16 * m + lr;
    float vr[8], vi[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const float2 v = k < 17 ? z(8 * lg + i, k) : make_float2(0.0f, 0.0f);
      vr[i] = v.x;
      vi[i] = v.y;
    }
    d.ar[m] = split8(vr, sz);
    d.ai[m] = split8(vi, sz);
  }
  return d;
}

// x = c2r(Z), handed to store(r, c0, x4) = x(r, c0 .. c0 + 3), unnormalised.
template <class X>
__device__ __forceinline__ void dft32_inv_compute(const DftSpec& d, float sz, const H8* T, int lane, X store) {
  const int lr = lane & 15, lg = lane >> 4;
  H2 br[2][2], bi[2][2];  // D tiles (spectral-column tile m, row tile n) of Y^T
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int n = 0; n < 2; n++) {
      const H8 fr = tile(T, DftTab::BR + n, lane), fi = tile(T, DftTab::BI + n, lane);
      const H8 fn = tile(T, DftTab::BN + n, lane);
      const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
      f4 yr = mma3(d.ar[m], fr, zero);
      yr = mma3(d.ai[m], fi, yr);
      f4 yi = mma3(d.ai[m], fr, zero);
      yi = mma3(d.ar[m], fn, yi);
      br[m][n] = split4(yr, 1.0f / 64.0f);  // |Y| <= 32 max|z|; the c2r weights add 2 x 17
      bi[m][n] = split4(yi, 1.0f / 64.0f);
    }
  const float back = 64.0f / sz;
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int n = 0; n < 2; n++) {
      f4 x = mma3(tile(T, DftTab::DR + m, lane), cat(br[0][n], br[1][n]), f4{0.0f, 0.0f, 0.0f, 0.0f});
      x = mma3(tile(T, DftTab::DN + m, lane), cat(bi[0][n], bi[1][n]), x);
      store(16 * n + lr, 16 * m + 4 * lg, f4{x[0] * back, x[1] * back, x[2] * back, x[3] * back});
    }
}

// Device copy of the constant tiles (built on the host in double, once per
// device); null with *err set on failure.
const H8* mfma_dft_table(hipError_t* err);

}  // namespace nf3d
