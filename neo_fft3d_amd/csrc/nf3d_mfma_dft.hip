// DFT-as-GEMM precision form, standalone: unnormalised 2D DFTs of 32x32 real
// blocks as dense DFT-matrix products on the gfx950 matrix cores (the block
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

// Forward: in [nb][32][32] float -> out [nb][32][17] float2 (FFTW r2c layout).
__global__ void __launch_bounds__(256) k_dft32_fwd(const float* in, float2* out, int nb, float sx, const H8* T) {
  const int lane = threadIdx.x & 63;
  const int blk = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (blk >= nb) return;  // whole wavefronts
  const float* X = in + (size_t)blk * 1024;
  const DftIn d = dft32_fwd_load([&](int r, int c) { return X[r * 32 + c]; }, sx, lane);
  float2* o = out + (size_t)blk * 544;
  dft32_fwd_compute(d, sx, T, lane, [&](int r, int k, float2 z) { o[r * 17 + k] = z; });
}

// Inverse: in [nb][32][17] float2 -> out [nb][32][32] float (FFTW c2r, unnormalised).
__global__ void __launch_bounds__(256) k_dft32_inv(const float2* in, float* out, int nb, float sz, const H8* T) {
  const int lane = threadIdx.x & 63;
  const int blk = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (blk >= nb) return;
  const float2* Z = in + (size_t)blk * 544;
  const DftSpec d = dft32_inv_load([&](int r, int k) { return Z[r * 17 + k]; }, sz, lane);
  float* o = out + (size_t)blk * 1024;
  dft32_inv_compute(d, sz, T, lane, [&](int r, int c0, f4 x) {
    *reinterpret_cast<float4*>(o + r * 32 + c0) = make_float4(x[0], x[1], x[2], x[3]);
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

// A tile (row tile m) of a 32x32 matrix M(row, col): lane l -> M[16m + l%16][K(l/16, i)]
template <class F, class K>
void a_tile(std::vector<H8>& T, int t, int m, F M, K kidx) {
  for (int l = 0; l < 64; l++) {
    double v[8];
    for (int i = 0; i < 8; i++) v[i] = M(16 * m + l % 16, kidx(l / 16, i));
    put(T, t, l, v);
  }
}

// B tile (column tile n): lane l -> M[K(l/16, i)][16n + l%16], memory K order
template <class F>
void b_tile(std::vector<H8>& T, int t, int n, F M) {
  for (int l = 0; l < 64; l++) {
    double v[8];
    for (int i = 0; i < 8; i++) v[i] = M(k_mem(l / 16, i), 16 * n + l % 16);
    put(T, t, l, v);
  }
}

std::vector<H8> build_tables() {
  const double w = 2.0 * 3.14159265358979323846 / 32.0;
  std::vector<H8> T((size_t)DftTab::COUNT * 64);
  auto C = [&](int j, int c) {
    if (c < 17) return std::cos(w * ((j * c) % 32));
    if (c >= 32 && c < 49) return -std::sin(w * ((j * (c - 32)) % 32));
    return 0.0;
  };
  auto Fr = [&](int a, int b) { return std::cos(w * ((a * b) % 32)); };
  auto Fi = [&](int a, int b) { return -std::sin(w * ((a * b) % 32)); };
  auto Fn = [&](int a, int b) { return std::sin(w * ((a * b) % 32)); };
  auto wk = [](int k) { return (k == 0 || k == 16) ? 1.0 : 2.0; };
  // Dr^T / Dn^T: row j (output column of x), column k (spectral column)
  auto DrT = [&](int j, int k) { return k < 17 ? wk(k) * std::cos(w * ((j * k) % 32)) : 0.0; };
  auto DnT = [&](int j, int k) { return k < 17 ? -wk(k) * std::sin(w * ((j * k) % 32)) : 0.0; };
  for (int n = 0; n < 4; n++) b_tile(T, DftTab::C + n, n, C);
  for (int m = 0; m < 2; m++) {
    a_tile(T, DftTab::FR + m, m, Fr, k_dtile);
    a_tile(T, DftTab::FI + m, m, Fi, k_dtile);
    a_tile(T, DftTab::DR + m, m, DrT, k_dtile);
    a_tile(T, DftTab::DN + m, m, DnT, k_dtile);
  }
  for (int n = 0; n < 2; n++) {
    b_tile(T, DftTab::BR + n, n, Fr);
    b_tile(T, DftTab::BI + n, n, Fi);
    b_tile(T, DftTab::BN + n, n, Fn);
  }
  return T;
}

}  // namespace

const H8* mfma_dft_table(hipError_t* err) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    *err = e;
    return nullptr;
  }
  // one copy of the tables per device, built on first use
  static std::mutex mu;
  static std::vector<H8*> tabs;
  std::lock_guard<std::mutex> lock(mu);
  if ((int)tabs.size() <= dev) tabs.resize(dev + 1, nullptr);
  if (!tabs[dev]) {
    std::vector<H8> h = build_tables();
    H8* d = nullptr;
    e = hipMalloc(&d, h.size() * sizeof(H8));
    if (e == hipSuccess) e = hipMemcpy(d, h.data(), h.size() * sizeof(H8), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      hipFree(d);
      (void)hipGetLastError();  // not left behind for a later launch check
      *err = e;
      return nullptr;
    }
    tabs[dev] = d;
  }
  *err = hipSuccess;
  return tabs[dev];
}

hipError_t launch_dft_mfma(const void* in, void* out, int nb, int n, int inverse, float scale, hipStream_t s) {
  if (n != 32 || nb < 0) return hipErrorInvalidValue;
  if (!(scale > 0.0f) || std::ldexp(1.0f, std::ilogb(scale)) != scale) return hipErrorInvalidValue;
  if (nb == 0) return hipSuccess;
  hipError_t e;
  const H8* T = mfma_dft_table(&e);
  if (!T) return e;
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
