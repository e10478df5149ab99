// DFT-as-GEMM precision form, standalone: unnormalised 2D DFTs of 32x32,
// 48x48 or 64x64 real blocks as dense DFT-matrix products on the gfx950 matrix cores (the block
// functions and the FP16 split are in nf3d_mfma.h), one wavefront per block,
// four blocks per workgroup; and the host build of the constant tiles, in
// double, cached per device.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <vector>

#include "nf3d_internal.h"
#include "nf3d_mfma.h"

namespace nf3d {

// Forward: in [nb][N][N] float -> out [nb][N][N/2+1] float2 (FFTW r2c layout).
template <int N>
__global__ void __launch_bounds__(256) k_dft_fwd(const float* in, float2* out, int nb, float sx, const H8* T) {
  constexpr int H = N / 2 + 1;
  const int lane = threadIdx.x & 63;
  const int blk = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (blk >= nb) return;  // whole wavefronts
  const float* X = in + (size_t)blk * N * N;
  const DftIn<N> d = dft_fwd_load<N>([&](int r, int c) { return X[r * N + c]; }, sx, lane);
  float2* o = out + (size_t)blk * N * H;
  dft_fwd_compute<N>(d, sx, T, lane, [&](int r, int k, float2 z) { o[r * H + k] = z; });
}

// Inverse: in [nb][N][N/2+1] float2 -> out [nb][N][N] float (FFTW c2r, unnormalised).
template <int N>
__global__ void __launch_bounds__(256) k_dft_inv(const float2* in, float* out, int nb, float sz, const H8* T) {
  constexpr int H = N / 2 + 1;
  const int lane = threadIdx.x & 63;
  const int blk = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (blk >= nb) return;
  const float2* Z = in + (size_t)blk * N * H;
  const DftSpec<N> d = dft_inv_load<N>([&](int r, int k) { return Z[r * H + k]; }, sz, lane);
  float* o = out + (size_t)blk * N * N;
  dft_inv_compute<N>(d, sz, T, lane, [&](int r, int c0, f4 x) {
    *reinterpret_cast<float4*>(o + r * N + c0) = make_float4(x[0], x[1], x[2], x[3]);
  });
}

// ----------------------------------------------------------------- tables
namespace {

void put(std::vector<H8>& T, int t, int lane, const double (&v)[8]) {
  H8 h;
  for (int i = 0; i < 8; i++) {
    const float x = (float)v[i];
    const _Float16 hi = (_Float16)x;
    h.hi[i] = hi;
    h.lo[i] = (_Float16)(x - (float)hi);
  }
  T[(size_t)t * 64 + lane] = h;
}

// K index of value i of lane group g: memory order, or the order in which a
// lane holds the rows of the previous product's two D tiles (see nf3d_mfma.h)
int k_mem(int g, int i) { return 8 * g + i; }
int k_dtile(int g, int i) { return i < 4 ? 4 * g + i : 16 + 4 * g + (i - 4); }

// A tile (row tile m, K step ks) of a matrix M(row, col):
// lane l -> M[16m + l%16][32ks + K(l/16, i)]
template <class F, class K>
void a_tile(std::vector<H8>& T, int t, int m, int ks, F M, K kidx) {
  for (int l = 0; l < 64; l++) {
    double v[8];
    for (int i = 0; i < 8; i++) v[i] = M(16 * m + l % 16, 32 * ks + kidx(l / 16, i));
    put(T, t, l, v);
  }
}

// B tile (K step ks, column tile n): lane l -> M[32ks + K(l/16, i)][16n + l%16], memory K order
template <class F>
void b_tile(std::vector<H8>& T, int t, int ks, int n, F M) {
  for (int l = 0; l < 64; l++) {
    double v[8];
    for (int i = 0; i < 8; i++) v[i] = M(32 * ks + k_mem(l / 16, i), 16 * n + l % 16);
    put(T, t, l, v);
  }
}

template <int N>
std::vector<H8> build_tables() {
  using S = DftShape<N>;
  const double w = 2.0 * 3.14159265358979323846 / N;
  std::vector<H8> T((size_t)S::COUNT * 64);
  const int H = S::H, half = 16 * S::NC;
  // indices past N (the K padding of N = 48) are zero entries
  auto C = [&](int j, int c) {
    if (j >= N) return 0.0;
    if (c < H) return std::cos(w * ((j * c) % N));
    if (c >= half && c - half < H) return -std::sin(w * ((j * (c - half)) % N));
    return 0.0;
  };
  auto Fr = [&](int a, int b) { return a < N && b < N ? std::cos(w * ((a * b) % N)) : 0.0; };
  auto Fi = [&](int a, int b) { return a < N && b < N ? -std::sin(w * ((a * b) % N)) : 0.0; };
  auto Fn = [&](int a, int b) { return a < N && b < N ? std::sin(w * ((a * b) % N)) : 0.0; };
  auto wk = [&](int k) { return (k == 0 || k == N / 2) ? 1.0 : 2.0; };
  // Dr^T / Dn^T: row j (output column of x), column k (spectral column; zero past H)
  auto DrT = [&](int j, int k) { return k < H ? wk(k) * std::cos(w * ((j * k) % N)) : 0.0; };
  auto DnT = [&](int j, int k) { return k < H ? -wk(k) * std::sin(w * ((j * k) % N)) : 0.0; };
  for (int ks = 0; ks < S::KS; ks++)
    for (int n = 0; n < 2 * S::NC; n++) b_tile(T, S::C + ks * 2 * S::NC + n, ks, n, C);
  for (int m = 0; m < S::R; m++) {
    for (int ks = 0; ks < S::KS; ks++) {
      a_tile(T, S::FR + m * S::KS + ks, m, ks, Fr, k_dtile);
      a_tile(T, S::FI + m * S::KS + ks, m, ks, Fi, k_dtile);
    }
    for (int ks = 0; ks < S::KSI; ks++) {
      a_tile(T, S::DR + m * S::KSI + ks, m, ks, DrT, k_dtile);
      a_tile(T, S::DN + m * S::KSI + ks, m, ks, DnT, k_dtile);
    }
  }
  for (int ks = 0; ks < S::KS; ks++)
    for (int n = 0; n < S::R; n++) {
      b_tile(T, S::BR + ks * S::R + n, ks, n, Fr);
      b_tile(T, S::BI + ks * S::R + n, ks, n, Fi);
      b_tile(T, S::BN + ks * S::R + n, ks, n, Fn);
    }
  return T;
}

}  // namespace

const H8* mfma_dft_table(int n, hipError_t* err) {
  if (n != 32 && n != 48 && n != 64) {
    *err = hipErrorInvalidValue;
    return nullptr;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    *err = e;
    return nullptr;
  }
  // one copy of the tables per device and block size, built on first use
  static std::mutex mu;
  static std::vector<H8*> tabs[3];
  std::lock_guard<std::mutex> lock(mu);
  std::vector<H8*>& tab = tabs[n == 32 ? 0 : n == 48 ? 1 : 2];
  if ((int)tab.size() <= dev) tab.resize(dev + 1, nullptr);
  if (!tab[dev]) {
    std::vector<H8> h = n == 32 ? build_tables<32>() : n == 48 ? build_tables<48>() : build_tables<64>();
    H8* d = nullptr;
    e = hipMalloc(&d, h.size() * sizeof(H8));
    if (e == hipSuccess) e = hipMemcpy(d, h.data(), h.size() * sizeof(H8), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      hipFree(d);
      (void)hipGetLastError();  // not left behind for a later launch check
      *err = e;
      return nullptr;
    }
    tab[dev] = d;
  }
  *err = hipSuccess;
  return tab[dev];
}

hipError_t launch_dft_mfma(const void* in, void* out, int nb, int n, int inverse, float scale, hipStream_t s) {
  if ((n != 32 && n != 48 && n != 64) || nb < 0) return hipErrorInvalidValue;
  if (nb > 0 && (!in || !out)) return hipErrorInvalidValue;
  if (!(scale > 0.0f) || std::ldexp(1.0f, std::ilogb(scale)) != scale) return hipErrorInvalidValue;
  if (nb == 0) return hipSuccess;
  hipError_t e;
  const H8* T = mfma_dft_table(n, &e);
  if (!T) return e;
  dim3 grid((nb + 3) / 4), block(256);
  const float2* zin = static_cast<const float2*>(in);
  const float* xin = static_cast<const float*>(in);
  if (n == 32) {
    if (inverse) hipLaunchKernelGGL(k_dft_inv<32>, grid, block, 0, s, zin, static_cast<float*>(out), nb, scale, T);
    else hipLaunchKernelGGL(k_dft_fwd<32>, grid, block, 0, s, xin, static_cast<float2*>(out), nb, scale, T);
  } else if (n == 48) {
    if (inverse) hipLaunchKernelGGL(k_dft_inv<48>, grid, block, 0, s, zin, static_cast<float*>(out), nb, scale, T);
    else hipLaunchKernelGGL(k_dft_fwd<48>, grid, block, 0, s, xin, static_cast<float2*>(out), nb, scale, T);
  } else {
    if (inverse) hipLaunchKernelGGL(k_dft_inv<64>, grid, block, 0, s, zin, static_cast<float*>(out), nb, scale, T);
    else hipLaunchKernelGGL(k_dft_fwd<64>, grid, block, 0, s, xin, static_cast<float2*>(out), nb, scale, T);
  }
  return hipGetLastError();
}

}  // namespace nf3d
