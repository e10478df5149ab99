// 32x32 block DFTs as dense DFT-matrix products on the gfx950 matrix cores,
// shared by the standalone transforms (nf3d_mfma_dft.hip) and the filter's
// per-frame kernel (k_tile with the matrix-core DFT).
//
// Every real operand is split into two FP16 halves, x = hi + lo with
// hi = fp16(x) and lo = fp16(x - hi), and a product is taken as
// hi*hi + hi*lo + lo*hi with FP32 accumulation (v_mfma_f32_16x16x16_f16, three
// per tile): about 22 significant bits per operand against FP32's 24, so the
// transforms land within a few FP32 ulps of an FFT.  Operands are scaled by
// powers of two (exact) to stay inside the FP16 range; the output is scaled
// back exactly.
//
// One wavefront owns one block.  Layouts of the 16x16x16 MFMA (lane l):
//   A (16x16, row-major M x K): lane l holds A[l % 16][4*(l/16) + i], i < 4
//   B (K x N):                  lane l holds B[4*(l/16) + i][l % 16]
//   D (M x N):                  lane l holds D[4*(l/16) + i][l % 16]
// so a D tile is directly the B operand of the next product.  The chains:
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
typedef float f4 __attribute__((ext_vector_type(4)));

struct H2 {
  half4 hi, lo;
};

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

__device__ __forceinline__ f4 mma3(const H2& a, const H2& b, f4 c) {
  c = __builtin_amdgcn_mfma_f32_16x16x16f16(a.lo, b.hi, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.lo, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.hi, c, 0, 0, 0);
  return c;
}

// Constant tiles, [tile][lane] of H2.
struct DftTab {
  static constexpr int C = 0;      // forward B tiles of C: (ks, n) -> ks * 4 + n, 8 tiles
  static constexpr int FR = 8;     // forward A tiles of Fr: (m, ks) -> m * 2 + ks, 4 tiles
  static constexpr int FI = 12;    // forward A tiles of Fi
  static constexpr int BR = 16;    // inverse B tiles of Fr: (ks, n) -> ks * 2 + n, 4 tiles
  static constexpr int BI = 20;    // inverse B tiles of Fi
  static constexpr int BN = 24;    // inverse B tiles of -Fi
  static constexpr int DR = 28;    // inverse A tiles of Dr^T: (m, ks) -> m * 2 + ks, 4 tiles
  static constexpr int DN = 32;    // inverse A tiles of Dn^T
  static constexpr int COUNT = 36;
};

__device__ __forceinline__ H2 tile(const H2* T, int t, int lane) { return T[t * 64 + lane]; }

// ---------------------------------------------------------------- forward
struct DftIn {
  H2 a[2][2];  // A tiles of x: (m, ks)
};

// x(r, c) -> operand tiles, times sx.
template <class X>
__device__ __forceinline__ DftIn dft32_fwd_load(X x, float sx, int lane) {
  const int lr = lane & 15, lg = lane >> 4;
  DftIn d;
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
      f4 v;
#pragma unroll
      for (int i = 0; i < 4; i++) v[i] = x(16 * m + lr, 16 * ks + 4 * lg + i);
      d.a[m][ks] = split4(v, sx);
    }
  return d;
}

// Z = F X C, handed to store(r, k, z) for k < 17 (each (r, k) once).
template <class Z>
__device__ __forceinline__ void dft32_fwd_compute(const DftIn& d, float sx, const H2* T, int lane, Z store) {
  const int lr = lane & 15, lg = lane >> 4;
  H2 yb[2][4];
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int n = 0; n < 4; n++) {
      f4 y = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < 2; ks++) y = mma3(d.a[m][ks], tile(T, DftTab::C + ks * 4 + n, lane), y);
      yb[m][n] = split4(y, 1.0f / 32.0f);  // |Y| <= 32 max|x|
    }
  const float back = 32.0f / sx;
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int n = 0; n < 2; n++) {
      f4 pr = {0, 0, 0, 0}, pi = {0, 0, 0, 0}, qr = {0, 0, 0, 0}, qi = {0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < 2; ks++) {
        const H2 fr = tile(T, DftTab::FR + m * 2 + ks, lane), fi = tile(T, DftTab::FI + m * 2 + ks, lane);
        pr = mma3(fr, yb[ks][n], pr);      // Fr Yr
        pi = mma3(fr, yb[ks][n + 2], pi);  // Fr Yi
        qr = mma3(fi, yb[ks][n], qr);      // Fi Yr
        qi = mma3(fi, yb[ks][n + 2], qi);  // Fi Yi
      }
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
  H2 ar[2][2], ai[2][2];  // A tiles of Z^T (real, imaginary): (m, ks)
};

// z(r, k) (k < 17) -> operand tiles, times sz.
template <class Z>
__device__ __forceinline__ DftSpec dft32_inv_load(Z z, float sz, int lane) {
  const int lr = lane & 15, lg = lane >> 4;
  DftSpec d;
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
      const int k = 16 * m + lr;
      f4 vr, vi;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const float2 v = k < 17 ? z(16 * ks + 4 * lg + i, k) : make_float2(0.0f, 0.0f);
        vr[i] = v.x;
        vi[i] = v.y;
      }
      d.ar[m][ks] = split4(vr, sz);
      d.ai[m][ks] = split4(vi, sz);
    }
  return d;
}

// x = c2r(Z), handed to store(r, c0, x4) = x(r, c0 .. c0 + 3), unnormalised.
template <class X>
__device__ __forceinline__ void dft32_inv_compute(const DftSpec& d, float sz, const H2* T, int lane, X store) {
  const int lr = lane & 15, lg = lane >> 4;
  H2 br[2][2], bi[2][2];  // B operands of the second product: [ks = m][n]
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int n = 0; n < 2; n++) {
      f4 yr = {0, 0, 0, 0}, yi = {0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < 2; ks++) {
        const H2 fr = tile(T, DftTab::BR + ks * 2 + n, lane), fi = tile(T, DftTab::BI + ks * 2 + n, lane);
        const H2 fn = tile(T, DftTab::BN + ks * 2 + n, lane);
        yr = mma3(d.ar[m][ks], fr, yr);
        yr = mma3(d.ai[m][ks], fi, yr);
        yi = mma3(d.ai[m][ks], fr, yi);
        yi = mma3(d.ar[m][ks], fn, yi);
      }
      br[m][n] = split4(yr, 1.0f / 64.0f);  // |Y| <= 32 max|z|; the c2r weights add 2 x 17
      bi[m][n] = split4(yi, 1.0f / 64.0f);
    }
  const float back = 64.0f / sz;
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int n = 0; n < 2; n++) {
      f4 x = {0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < 2; ks++) {
        x = mma3(tile(T, DftTab::DR + m * 2 + ks, lane), br[ks][n], x);
        x = mma3(tile(T, DftTab::DN + m * 2 + ks, lane), bi[ks][n], x);
      }
      store(16 * n + lr, 16 * m + 4 * lg, f4{x[0] * back, x[1] * back, x[2] * back, x[3] * back});
    }
}

// Device copy of the constant tiles (built on the host in double, once per
// device); null with *err set on failure.
const H2* mfma_dft_table(hipError_t* err);

}  // namespace nf3d
