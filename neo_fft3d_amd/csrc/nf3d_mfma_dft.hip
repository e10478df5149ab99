// DFT-as-GEMM precision form: unnormalised 2D DFTs of 32x32 real blocks as
// dense DFT-matrix products on the gfx950 matrix cores.
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
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <vector>

#include "nf3d_internal.h"

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

// Constant tiles, [tile][lane] of H2, in the order of DftTab.
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

// Forward: in [nb][32][32] float -> out [nb][32][17] float2 (FFTW r2c layout).
__global__ void __launch_bounds__(256) k_dft32_fwd(const float* in, float2* out, int nb, float sx, const H2* T) {
  const int lane = threadIdx.x & 63;
  const int blk = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (blk >= nb) return;  // whole wavefronts
  const int lr = lane & 15, lg = lane >> 4;
  const float* X = in + (size_t)blk * 1024;
  H2 xa[2][2];
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
      const float4 v = *reinterpret_cast<const float4*>(X + (16 * m + lr) * 32 + 16 * ks + 4 * lg);
      xa[m][ks] = split4(f4{v.x, v.y, v.z, v.w}, sx);
    }
  // Y = X C: rows 32 (m), columns 64 (n: Yr in tiles 0-1, Yi in tiles 2-3)
  H2 yb[2][4];
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int n = 0; n < 4; n++) {
      f4 y = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < 2; ks++) y = mma3(xa[m][ks], tile(T, DftTab::C + ks * 4 + n, lane), y);
      yb[m][n] = split4(y, 1.0f / 32.0f);  // |Y| <= 32 max|x|
    }
  // Z = F Y
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
        float2* o = out + (size_t)blk * 544 + k;
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int r = 16 * m + 4 * lg + i;
          o[r * 17] = make_float2((pr[i] - qi[i]) * back, (pi[i] + qr[i]) * back);
        }
      }
    }
}

// Inverse: in [nb][32][17] float2 -> out [nb][32][32] float (FFTW c2r, unnormalised).
__global__ void __launch_bounds__(256) k_dft32_inv(const float2* in, float* out, int nb, float sz, const H2* T) {
  const int lane = threadIdx.x & 63;
  const int blk = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (blk >= nb) return;
  const int lr = lane & 15, lg = lane >> 4;
  const float2* Z = in + (size_t)blk * 544;
  // A tiles of Z^T: rows k = 16 m + lr (< 17), columns r = 16 ks + 4 lg + i
  H2 ar[2][2], ai[2][2];
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
      const int k = 16 * m + lr;
      f4 vr, vi;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const float2 z = k < 17 ? Z[(16 * ks + 4 * lg + i) * 17 + k] : make_float2(0.0f, 0.0f);
        vr[i] = z.x;
        vi[i] = z.y;
      }
      ar[m][ks] = split4(vr, sz);
      ai[m][ks] = split4(vi, sz);
    }
  // Y^T = Z^T conj(F): Y^T_r = Zr Fr + Zi Fi, Y^T_i = Zi Fr + Zr (-Fi)
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
        yr = mma3(ar[m][ks], fr, yr);
        yr = mma3(ai[m][ks], fi, yr);
        yi = mma3(ai[m][ks], fr, yi);
        yi = mma3(ar[m][ks], fn, yi);
      }
      br[m][n] = split4(yr, 1.0f / 64.0f);  // |Y| <= 32 max|z|; the c2r weights add 2 x 17
      bi[m][n] = split4(yi, 1.0f / 64.0f);
    }
  // x^T = Dr^T Y^T_r + Dn^T Y^T_i: rows j (m), columns r' (n)
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
      const int r = 16 * n + lr;
      *reinterpret_cast<float4*>(out + (size_t)blk * 1024 + r * 32 + 16 * m + 4 * lg) =
          make_float4(x[0] * back, x[1] * back, x[2] * back, x[3] * back);
    }
}

// ----------------------------------------------------------------- tables
namespace {

void put(std::vector<H2>& T, int t, int lane, const double (&v)[4]) {
  H2 h;
  for (int i = 0; i < 4; i++) {
    const float x = (float)v[i];
    const _Float16 hi = (_Float16)x;
    h.hi[i] = hi;
    h.lo[i] = (_Float16)(x - (float)hi);
  }
  T[(size_t)t * 64 + lane] = h;
}

// A tile (m, ks) of a 32x32 matrix M(row, col): lane l -> M[16m + l%16][16ks + 4(l/16) + i]
template <class F>
void a_tile(std::vector<H2>& T, int t, int m, int ks, F M) {
  for (int l = 0; l < 64; l++) {
    double v[4];
    for (int i = 0; i < 4; i++) v[i] = M(16 * m + l % 16, 16 * ks + 4 * (l / 16) + i);
    put(T, t, l, v);
  }
}

// B tile (ks, n): lane l -> M[16ks + 4(l/16) + i][16n + l%16]
template <class F>
void b_tile(std::vector<H2>& T, int t, int ks, int n, F M) {
  for (int l = 0; l < 64; l++) {
    double v[4];
    for (int i = 0; i < 4; i++) v[i] = M(16 * ks + 4 * (l / 16) + i, 16 * n + l % 16);
    put(T, t, l, v);
  }
}

std::vector<H2> build_tables() {
  const double w = 2.0 * 3.14159265358979323846 / 32.0;
  std::vector<H2> T((size_t)DftTab::COUNT * 64);
  auto C = [&](int j, int c) {
    if (c < 17) return std::cos(w * ((j * c) % 32));
    if (c >= 32 && c < 49) return -std::sin(w * ((j * (c - 32)) % 32));
    return 0.0;
  };
  auto Fr = [&](int a, int b) { return std::cos(w * ((a * b) % 32)); };
  auto Fi = [&](int a, int b) { return -std::sin(w * ((a * b) % 32)); };
  auto Fn = [&](int a, int b) { return std::sin(w * ((a * b) % 32)); };
  auto wk = [](int k) { return (k == 0 || k == 16) ? 1.0 : 2.0; };
  auto Dr = [&](int j, int k) { return k < 17 ? wk(k) * std::cos(w * ((j * k) % 32)) : 0.0; };
  auto Dn = [&](int j, int k) { return k < 17 ? -wk(k) * std::sin(w * ((j * k) % 32)) : 0.0; };
  for (int ks = 0; ks < 2; ks++)
    for (int n = 0; n < 4; n++) b_tile(T, DftTab::C + ks * 4 + n, ks, n, C);
  for (int m = 0; m < 2; m++)
    for (int ks = 0; ks < 2; ks++) {
      a_tile(T, DftTab::FR + m * 2 + ks, m, ks, Fr);
      a_tile(T, DftTab::FI + m * 2 + ks, m, ks, Fi);
      a_tile(T, DftTab::DR + m * 2 + ks, m, ks, Dr);
      a_tile(T, DftTab::DN + m * 2 + ks, m, ks, Dn);
    }
  for (int ks = 0; ks < 2; ks++)
    for (int n = 0; n < 2; n++) {
      b_tile(T, DftTab::BR + ks * 2 + n, ks, n, Fr);
      b_tile(T, DftTab::BI + ks * 2 + n, ks, n, Fi);
      b_tile(T, DftTab::BN + ks * 2 + n, ks, n, Fn);
    }
  return T;
}

}  // namespace

hipError_t launch_dft_mfma(const void* in, void* out, int nb, int n, int inverse, float scale, hipStream_t s) {
  if (n != 32 || nb < 0) return hipErrorInvalidValue;
  if (!(scale > 0.0f) || std::ldexp(1.0f, std::ilogb(scale)) != scale) return hipErrorInvalidValue;
  if (nb == 0) return hipSuccess;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  // one copy of the tables per device, built on first use
  static std::mutex mu;
  static std::vector<H2*> tabs;
  H2* T = nullptr;
  {
    std::lock_guard<std::mutex> lock(mu);
    if ((int)tabs.size() <= dev) tabs.resize(dev + 1, nullptr);
    if (!tabs[dev]) {
      std::vector<H2> h = build_tables();
      e = hipMalloc(&tabs[dev], h.size() * sizeof(H2));
      if (e != hipSuccess) return e;
      e = hipMemcpy(tabs[dev], h.data(), h.size() * sizeof(H2), hipMemcpyHostToDevice);
      if (e != hipSuccess) return e;
    }
    T = tabs[dev];
  }
  dim3 grid((nb + 3) / 4), block(256);
  if (inverse)
    hipLaunchKernelGGL(k_dft32_inv, grid, block, 0, s, static_cast<const float2*>(in), static_cast<float*>(out), nb,
                       scale, T);
  else
    hipLaunchKernelGGL(k_dft32_fwd, grid, block, 0, s, static_cast<const float*>(in), static_cast<float2*>(out), nb,
                       scale, T);
  return hipGetLastError();
}

}  // namespace nf3d
