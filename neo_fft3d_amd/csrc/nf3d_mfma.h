// N x N block DFTs (N = 32, 48 or 64) as dense DFT-matrix products on the gfx950
// matrix cores, shared by the standalone transforms (nf3d_mfma_dft.hip) and
// the filter's kernels (k_run / k_tile with the matrix-core DFT).
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
// A K of 64 is two products of 32; a K of 48 is padded to 64 with zero
// operands (zero table entries and, for the products of two D-tile chains,
// a zero fourth row tile).  With H = N/2 + 1 kept frequencies in
// NC = ceil(H / 16) column tiles, the chains are:
//   forward  Y = X C        (C: N x 2*16*NC, columns k < H cos, 16*NC + k -sin)
//            Z = F Y        (F = Fr + i Fi, Fr = cos, Fi = -sin; N x N)
//            as P = Fr [Yr|Yi], Q = Fi [Yr|Yi]: Zr = P_r - Q_i, Zi = P_i + Q_r
//   inverse  Y^T = Z^T conj(F)            (per spectral column, all rows)
//            x^T = Dr^T Y^T_r + Dn^T Y^T_i (Dr = w cos, Dn = -w sin, w = 1 at
//            k = 0 and N/2, else 2: the c2r of the last dimension, the
//            imaginary parts of columns 0 and N/2 dropped as FFTW's c2r does;
//            K = the spectral columns, zero past H up to a multiple of 32)
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

// Tile counts of an N x N block.
template <int N>
struct DftShape {
  static_assert(N == 32 || N == 48 || N == 64, "matrix-core DFT: 32x32, 48x48 or 64x64 blocks");
  static constexpr int R = N / 16;              // 16-row tiles of the block
  static constexpr int KS = (N + 31) / 32;      // K steps over the block's rows / columns (zero past N)
  static constexpr int H = N / 2 + 1;           // kept frequencies
  // power of two >= N: intermediate products are scaled by 1/P2 before they
  // are split, so the scaling is exact (a scale like 1/48 rounds the product
  // once to FP32 for the low half, while the compiler may round the high
  // half straight from the exact product: near an FP16 tie the two halves
  // then disagree by one FP16 step)
  static constexpr int P2 = N <= 32 ? 32 : 64;
  static constexpr int NC = (H + 15) / 16;      // 16-column tiles of the kept frequencies
  static constexpr int KSI = (16 * NC + 31) / 32;  // K steps over the spectral columns (inverse)
  // constant tiles, [tile][lane] of H8
  static constexpr int C = 0;                   // forward B tiles of C (memory K order): (ks, n < 2 NC)
  static constexpr int FR = C + KS * 2 * NC;    // forward A tiles of Fr (D-tile K order): (m < R, ks)
  static constexpr int FI = FR + R * KS;        // forward A tiles of Fi
  static constexpr int BR = FI + R * KS;        // inverse B tiles of Fr (memory K order): (ks, n < R)
  static constexpr int BI = BR + KS * R;        // inverse B tiles of Fi
  static constexpr int BN = BI + KS * R;        // inverse B tiles of -Fi
  static constexpr int DR = BN + KS * R;        // inverse A tiles of Dr^T (D-tile K order): (m < R, ks < KSI)
  static constexpr int DN = DR + R * KSI;       // inverse A tiles of Dn^T
  static constexpr int COUNT = DN + R * KSI;
};

__device__ __forceinline__ H8 tile(const H8* T, int t, int lane) { return T[t * 64 + lane]; }

__device__ __forceinline__ H8 split8(const float (&v)[8], float s) {
  f4 v0 = {v[0], v[1], v[2], v[3]}, v1 = {v[4], v[5], v[6], v[7]};
  return cat(split4(v0, s), split4(v1, s));
}

// ---------------------------------------------------------------- forward
template <int N>
struct DftIn {
  H8 a[DftShape<N>::R][DftShape<N>::KS];  // A tiles of x: row tile m, K step ks
};

// x(r, c) -> operand tiles, times sx.
template <int N, class X>
__device__ __forceinline__ DftIn<N> dft_fwd_load(X x, float sx, int lane) {
  using S = DftShape<N>;
  const int lr = lane & 15, lg = lane >> 4;
  DftIn<N> d;
#pragma unroll
  for (int m = 0; m < S::R; m++)
#pragma unroll
    for (int ks = 0; ks < S::KS; ks++) {
      float v[8];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int c = 32 * ks + 8 * lg + i;
        if constexpr (N % 32 == 0) v[i] = x(16 * m + lr, c);
        else v[i] = c < N ? x(16 * m + lr, c) : 0.0f;  // K padding
      }
      d.a[m][ks] = split8(v, sx);
    }
  return d;
}

// Z = F X C, handed to store(r, k, z) for k < N/2 + 1 (each (r, k) once).
// Column tile by column tile: the Y tiles of one spectral column tile (its
// cos and sin halves, every row tile) are formed, then used by the second
// product for every output row tile, so only those tiles are live.
template <int N, class Z>
__device__ __forceinline__ void dft_fwd_compute(const DftIn<N>& d, float sx, const H8* T, int lane, Z store) {
  using S = DftShape<N>;
  const int lr = lane & 15, lg = lane >> 4;
  const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
  const float back = (float)S::P2 / sx;
#pragma unroll
  for (int n = 0; n < S::NC; n++) {
    H2 yr[2 * S::KS], yi[2 * S::KS];  // row tiles of Y; past R (K padding) zero
#pragma unroll
    for (int m = S::R; m < 2 * S::KS; m++) {
      yr[m] = split4(zero, 1.0f);
      yi[m] = yr[m];
    }
#pragma unroll
    for (int m = 0; m < S::R; m++) {
      f4 a = zero, b = zero;
#pragma unroll
      for (int ks = 0; ks < S::KS; ks++) {
        a = mma3(d.a[m][ks], tile(T, S::C + ks * 2 * S::NC + n, lane), a);
        b = mma3(d.a[m][ks], tile(T, S::C + ks * 2 * S::NC + n + S::NC, lane), b);
      }
      yr[m] = split4(a, 1.0f / S::P2);  // |Y| <= N max|x|
      yi[m] = split4(b, 1.0f / S::P2);
    }
#pragma unroll
    for (int m = 0; m < S::R; m++) {
      f4 pr = zero, pi = zero, qr = zero, qi = zero;
#pragma unroll
      for (int ks = 0; ks < S::KS; ks++) {
        const H8 fr = tile(T, S::FR + m * S::KS + ks, lane), fi = tile(T, S::FI + m * S::KS + ks, lane);
        const H8 br = cat(yr[2 * ks], yr[2 * ks + 1]), bi = cat(yi[2 * ks], yi[2 * ks + 1]);
        pr = mma3(fr, br, pr);  // Fr Yr
        pi = mma3(fr, bi, pi);  // Fr Yi
        qr = mma3(fi, br, qr);  // Fi Yr
        qi = mma3(fi, bi, qi);  // Fi Yi
      }
      const int k = 16 * n + lr;
      if (k < S::H) {
#pragma unroll
        for (int i = 0; i < 4; i++)
          store(16 * m + 4 * lg + i, k, make_float2((pr[i] - qi[i]) * back, (pi[i] + qr[i]) * back));
      }
    }
  }
}

// ---------------------------------------------------------------- inverse
template <int N>
struct DftSpec {
  // A tiles of Z^T (real, imaginary): spectral-column tile m, K step ks over the rows
  H8 ar[DftShape<N>::NC][DftShape<N>::KS], ai[DftShape<N>::NC][DftShape<N>::KS];
};

// z(r, k) (k < N/2 + 1) -> operand tiles, times sz.
template <int N, class Z>
__device__ __forceinline__ DftSpec<N> dft_inv_load(Z z, float sz, int lane) {
  using S = DftShape<N>;
  const int lr = lane & 15, lg = lane >> 4;
  DftSpec<N> d;
#pragma unroll
  for (int m = 0; m < S::NC; m++)
#pragma unroll
    for (int ks = 0; ks < S::KS; ks++) {
      const int k = 16 * m + lr;
      float vr[8], vi[8];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int r = 32 * ks + 8 * lg + i;
        const float2 v = k < S::H && (N % 32 == 0 || r < N) ? z(r, k) : make_float2(0.0f, 0.0f);
        vr[i] = v.x;
        vi[i] = v.y;
      }
      d.ar[m][ks] = split8(vr, sz);
      d.ai[m][ks] = split8(vi, sz);
    }
  return d;
}

// x = c2r(Z), handed to store(r, c0, x4) = x(r, c0 .. c0 + 3), unnormalised.
// Row tile by row tile of x: the Y^T tiles of one row tile (every spectral
// column tile) are formed, then used by the second product for every output
// column tile, so only those tiles are live.
template <int N, class X>
__device__ __forceinline__ void dft_inv_compute(const DftSpec<N>& d, float sz, const H8* T, int lane, X store) {
  using S = DftShape<N>;
  const int lr = lane & 15, lg = lane >> 4;
  const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
  const float back = (float)(2 * S::P2) / sz;
#pragma unroll
  for (int n = 0; n < S::R; n++) {
    H2 br[2 * S::KSI], bi[2 * S::KSI];  // D tiles (spectral-column tile m, this row tile) of Y^T
#pragma unroll
    for (int m = 0; m < 2 * S::KSI; m++) {
      if (m >= S::NC) {  // spectral columns past the kept ones: zero operands
        br[m] = split4(zero, 1.0f);
        bi[m] = br[m];
        continue;
      }
      f4 yr = zero, yi = zero;
#pragma unroll
      for (int ks = 0; ks < S::KS; ks++) {
        const H8 fr = tile(T, S::BR + ks * S::R + n, lane), fi = tile(T, S::BI + ks * S::R + n, lane);
        const H8 fn = tile(T, S::BN + ks * S::R + n, lane);
        yr = mma3(d.ar[m][ks], fr, yr);
        yr = mma3(d.ai[m][ks], fi, yr);
        yi = mma3(d.ai[m][ks], fr, yi);
        yi = mma3(d.ar[m][ks], fn, yi);
      }
      br[m] = split4(yr, 1.0f / (2 * S::P2));  // |Y| <= N max|z|; the c2r weights add 2 x (N/2 + 1)
      bi[m] = split4(yi, 1.0f / (2 * S::P2));
    }
#pragma unroll
    for (int m = 0; m < S::R; m++) {
      f4 x = zero;
#pragma unroll
      for (int ks = 0; ks < S::KSI; ks++) {
        x = mma3(tile(T, S::DR + m * S::KSI + ks, lane), cat(br[2 * ks], br[2 * ks + 1]), x);
        x = mma3(tile(T, S::DN + m * S::KSI + ks, lane), cat(bi[2 * ks], bi[2 * ks + 1]), x);
      }
      store(16 * n + lr, 16 * m + 4 * lg, f4{x[0] * back, x[1] * back, x[2] * back, x[3] * back});
    }
  }
}

// Device copy of the constant tiles of N x N blocks (built on the host in
// double, once per device and N); null with *err set on failure.
const H8* mfma_dft_table(int n, hipError_t* err);

}  // namespace nf3d
