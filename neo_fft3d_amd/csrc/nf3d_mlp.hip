// Learned noise PSD: a small MLP over a block's log-power spectrum
// (n_in = bh * (bw/2+1) inputs -> hidden -> hidden -> n_out log-PSD outputs,
// GELU), trained and run on the matrix cores.
//
// Numerics: BF16 operands (weights, activations, gradients) with FP32
// accumulation (v_mfma_f32_16x16x32_bf16); FP32 master weights, biases,
// pre-activations and Adam moments.
//
// Kernels:
//   k_gemm<TA, TB, EPI, TM, TN>
//                         C(m, n) = sum_k A(m, k) B(k, n) over a TM x TN tile
//                         per workgroup (4 waves of TM/2 x TN/2, 16x16 MFMA
//                         tiles; 64x64, or 32x32 when the 64x64 grid would
//                         leave CUs idle), K in steps of 64 staged through
//                         LDS, next step's operands loaded into registers
//                         during the MFMAs;
//                         epilogues: bias + GELU, bias, bias + exp (PSD),
//                         GELU backward, plain FP32 (weight gradients)
//   k_fwdp / k_fwdw / k_fwd
//                         inference layers (K <= 256, <= 1280, larger):
//                         fragments loaded straight from memory against the
//                         transposed weights, coalesced epilogue
//   k_transpose           the transposed BF16 weights, after each update
//   k_features_rows       log-power features of compact block spectra
//   k_whittle             Whittle loss y + t e^-y and its gradient
//   k_colsum              bias gradients (column sums, two passes)
//   k_adam                Adam update of the FP32 master and its BF16 copy
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/nf3d.h"
#include "nf3d_mlp_batch.h"

namespace nf3d_mlp_k {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TK = 64, NT = 256;
constexpr int LDK = TK + 8;  // LDS row: 64 k-values + 8 pad (144 bytes, 16-byte aligned)

enum Epi : int {
  EPI_F32 = 0,        // C = acc
  EPI_BIAS_GELU = 1,  // z = acc + bias -> C = z (if C), H = gelu(z)
  EPI_BIAS = 2,       // C = acc + bias
  EPI_BIAS_EXP = 3,   // C = exp(acc + bias) * scale
  EPI_GELU_BWD = 4,   // g = acc * gelu'(Zp) -> H = g, C = g (if C)
};

struct GemmArgs {
  const __bf16* A;  // A(m, k) = TA ? A[k * lda + m] : A[m * lda + k]
  const __bf16* B;  // B(k, n) = TB ? B[n * ldb + k] : B[k * ldb + n]
  long long lda, ldb;
  int M, N, K;
  const float* bias;  // [N]
  float* C;
  long long ldc;
  __bf16* H;
  long long ldh;
  const float* Zp;  // pre-activation of the GELU being differentiated
  long long ldz;
  float scale;
  // split-K (EPI_F32 only): workgroup z sums k in [z*kc, min(K, (z+1)*kc))
  // into C + z*cz; kc = K and cz = 0 without a split
  int kc;
  long long cz;
};

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// One operand tile (R rows of the outer index x 64 k) as R*8 chunks of 8
// contiguous elements; each lane loads R*8/256 of them.  KC: k is the
// contiguous index in memory (row-major A, transposed B), else the outer
// index is.
template <bool KC, int R>
struct TileLoad {
  static constexpr int NCH = R * 8 / NT;
  static constexpr int OC = R / 8;  // chunks per k-row when the outer index is contiguous
  uint4 v[NCH];
  __device__ void load(const __bf16* P, long long ld, int o0, int olim, int k0, int klim, int t) {
#pragma unroll
    for (int h = 0; h < NCH; h++) {
      const int c = t + h * NT;
      int o, k;
      if constexpr (KC) {
        o = c >> 3;
        k = (c & 7) * 8;
      } else {
        k = c / OC;
        o = (c % OC) * 8;
      }
      const int go = o0 + o, gk = k0 + k;
      v[h] = make_uint4(0, 0, 0, 0);
      if (go < olim && gk < klim)
        v[h] = *reinterpret_cast<const uint4*>(KC ? P + (long long)go * ld + gk : P + (long long)gk * ld + go);
    }
  }
  __device__ void store(__bf16* S, int t) const {
#pragma unroll
    for (int h = 0; h < NCH; h++) {
      const int c = t + h * NT;
      if constexpr (KC) {
        *reinterpret_cast<uint4*>(S + (c >> 3) * LDK + (c & 7) * 8) = v[h];
      } else {
        const int k = c / OC, o = (c % OC) * 8;
        __bf16 e[8];
        memcpy(e, &v[h], 16);
#pragma unroll
        for (int i = 0; i < 8; i++) S[(o + i) * LDK + k] = e[i];
      }
    }
  }
};

template <bool TA, bool TB, int EPI, int TM, int TN>
__global__ void __launch_bounds__(NT) k_gemm(GemmArgs g) {
  constexpr int MI = TM / 32, NJ = TN / 32;  // 16x16 MFMA tiles per wave
  __shared__ __attribute__((aligned(16))) __bf16 As[TM * LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[TN * LDK];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int wm = (w & 1) * (TM / 2), wn = (w >> 1) * (TN / 2);
  const int kb = blockIdx.z * g.kc, ke = min(g.K, kb + g.kc);
  TileLoad<!TA, TM> la;
  TileLoad<TB, TN> lb;
  la.load(g.A, g.lda, m0, g.M, kb, ke, t);
  lb.load(g.B, g.ldb, n0, g.N, kb, ke, t);
  f32x4 acc[MI][NJ];
#pragma unroll
  for (int i = 0; i < MI; i++)
#pragma unroll
    for (int j = 0; j < NJ; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const int r16 = lane & 15, kg = (lane >> 4) * 8;
  for (int k0 = kb; k0 < ke; k0 += TK) {
    __syncthreads();  // the previous step's MFMAs are done with the tiles
    la.store(As, t);
    lb.store(Bs, t);
    __syncthreads();
    if (k0 + TK < ke) {
      la.load(g.A, g.lda, m0, g.M, k0 + TK, ke, t);
      lb.load(g.B, g.ldb, n0, g.N, k0 + TK, ke, t);
    }
#pragma unroll
    for (int kk = 0; kk < TK; kk += 32) {
      bf16x8 a[MI], b[NJ];
#pragma unroll
      for (int i = 0; i < MI; i++) a[i] = *reinterpret_cast<const bf16x8*>(As + (wm + i * 16 + r16) * LDK + kk + kg);
#pragma unroll
      for (int j = 0; j < NJ; j++) b[j] = *reinterpret_cast<const bf16x8*>(Bs + (wn + j * 16 + r16) * LDK + kk + kg);
#pragma unroll
      for (int i = 0; i < MI; i++)
#pragma unroll
        for (int j = 0; j < NJ; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
  // epilogue: lane holds rows 4*(lane/16) + r of each 16x16 tile, column lane % 16
#pragma unroll
  for (int j = 0; j < NJ; j++) {
    const int n = n0 + wn + j * 16 + r16;
    if (n >= g.N) continue;
    const float bias = (EPI == EPI_BIAS_GELU || EPI == EPI_BIAS || EPI == EPI_BIAS_EXP) ? g.bias[n] : 0.0f;
#pragma unroll
    for (int i = 0; i < MI; i++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
        if (m >= g.M) continue;
        const float v = acc[i][j][r];
        if constexpr (EPI == EPI_F32) {
          g.C[blockIdx.z * g.cz + (long long)m * g.ldc + n] = v;
        } else if constexpr (EPI == EPI_BIAS_GELU) {
          const float z = v + bias;
          if (g.C) g.C[(long long)m * g.ldc + n] = z;
          g.H[(long long)m * g.ldh + n] = (__bf16)gelu(z);
        } else if constexpr (EPI == EPI_BIAS) {
          g.C[(long long)m * g.ldc + n] = v + bias;
        } else if constexpr (EPI == EPI_BIAS_EXP) {
          const float z = fminf(v + bias, 80.0f);
          g.C[(long long)m * g.ldc + n] = expf(z) * g.scale;
        } else {
          const float d = v * gelu_grad(g.Zp[(long long)m * g.ldz + n]);
          g.H[(long long)m * g.ldh + n] = (__bf16)d;
          if (g.C) g.C[(long long)m * g.ldc + n] = d;
        }
      }
  }
}

// One workgroup per row (block) of a batch, four coefficients per lane, the
// row's compact spectrum taken from its segment (up to 8 arrays of
// consecutive rows, e.g. the planes of a frame):
// x = (ln(|c|^2 + eps) - mu[k]) * isd[k].
struct FeatSegs {
  const float2* spec[8];
  int row0[9];  // segment s holds rows [row0[s], row0[s + 1])
  int n;
};
__global__ void k_features_rows(FeatSegs sg, __bf16* X, int d, const float* mu, const float* isd, float eps) {
  const int r = blockIdx.x, k = threadIdx.x * 4;
  if (k >= d) return;
  int s = 0;
  while (s + 1 < sg.n && r >= sg.row0[s + 1]) s++;
  const float2* src = sg.spec[s] + (long long)(r - sg.row0[s]) * d + k;
  const float4 a = *reinterpret_cast<const float4*>(src), b = *reinterpret_cast<const float4*>(src + 2);
  const float4 mv = *reinterpret_cast<const float4*>(mu + k), iv = *reinterpret_cast<const float4*>(isd + k);
  const float p0 = a.x * a.x + a.y * a.y + eps, p1 = a.z * a.z + a.w * a.w + eps;
  const float p2 = b.x * b.x + b.y * b.y + eps, p3 = b.z * b.z + b.w * b.w + eps;
  __bf16 o[4] = {(__bf16)((logf(p0) - mv.x) * iv.x), (__bf16)((logf(p1) - mv.y) * iv.y),
                 (__bf16)((logf(p2) - mv.z) * iv.z), (__bf16)((logf(p3) - mv.w) * iv.w)};
  uint2 v;
  memcpy(&v, o, 8);
  *reinterpret_cast<uint2*>(X + (long long)r * d + k) = v;
}

// Rows idx[r] of src into row r of dst (rows of `words` 16-byte words).
__global__ void k_gather_rows(const uint4* src, const int* idx, uint4* dst, int rows, int words) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)rows * words) return;
  const int r = (int)(i / words), w = (int)(i - (long long)r * words);
  dst[i] = src[(long long)idx[r] * words + w];
}

__global__ void k_to_bf16(const float* x, __bf16* y, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) y[i] = (__bf16)x[i];
}

__global__ void k_periodogram(const float2* spec, float* T, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) {
    const float2 c = spec[i];
    T[i] = c.x * c.x + c.y * c.y;
  }
}

// Whittle negative log-likelihood of periodogram values t under the
// predicted log-PSD y: mean of y + t e^-y; gradient (1 - t e^-y) / count.
// Per-workgroup partial sums of the loss (the host adds them).
__global__ void __launch_bounds__(256) k_whittle(const float* Y, const float* T, __bf16* dY, float* dYf,
                                                 long long total, float inv_count, float* partial) {
  __shared__ float red[256];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < total) {
    const float y = fminf(fmaxf(Y[i], -80.0f), 80.0f);
    const float te = T[i] * expf(-y);
    l = y + te;
    const float gr = (1.0f - te) * inv_count;
    dY[i] = (__bf16)gr;
    dYf[i] = gr;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Column sums of X [rows][cols]: pass 1 writes part[slice][col] over row
// slices of CS_ROWS rows (a workgroup: 64 columns x 4 row groups, each lane
// with its CS_ROWS/4 loads in flight together), pass 2 adds the slices in
// order (deterministic).  With 32-row slices a 4096-row batch gives 128
// slices, so even 256 columns fill 512 workgroups.
constexpr int CS_ROWS = 32;
__global__ void __launch_bounds__(256) k_colsum1(const float* X, int rows, int cols, float* part) {
  __shared__ float red[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), q = threadIdx.x >> 6;
  const int r0 = blockIdx.y * CS_ROWS;
  float v[CS_ROWS / 4];
#pragma unroll
  for (int u = 0; u < CS_ROWS / 4; u++) {
    const int r = r0 + q + 4 * u;
    v[u] = (c < cols && r < rows) ? X[(long long)r * cols + c] : 0.0f;
  }
  float s = 0.0f;
#pragma unroll
  for (int u = 0; u < CS_ROWS / 4; u++) s += v[u];
  red[q][threadIdx.x & 63] = s;
  __syncthreads();
  if (q == 0 && c < cols)
    part[blockIdx.y * cols + c] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}
__global__ void k_colsum2(const float* part, int cols, int slices, float* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  // the slices added in order, their loads sixteen at a time in flight (a
  // few hundred lanes in all: one load round trip per slice otherwise)
  float s = 0.0f;
  int q = 0;
  for (; q + 16 <= slices; q += 16) {
    float v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = part[(q + u) * cols + c];
#pragma unroll
    for (int u = 0; u < 16; u++) s += v[u];
  }
  for (; q < slices; q++) s += part[q * cols + c];
  out[c] = s;
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, __bf16* pb, long long n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  p[i] = pi;
  pb[i] = (__bf16)pi;
}

inline unsigned blocks(long long n, int t) { return (unsigned)((n + t - 1) / t); }

// out[i] = the S split-K partials of element i added in split order.
__global__ void k_splitsum(const float* part, int S, long long count, float* out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float v = part[i];
  for (int z = 1; z < S; z++) v += part[z * count + i];
  out[i] = v;
}

// 64x64 tiles, or 32x32 when the 64x64 grid has fewer than two workgroups
// per CU (256 CUs).
template <bool TA, bool TB, int EPI>
hipError_t gemm(const GemmArgs& g, hipStream_t s) {
  const long long big = (long long)((g.N + 63) / 64) * ((g.M + 63) / 64);
  if (big >= 512) {
    dim3 grid((g.N + 63) / 64, (g.M + 63) / 64);
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI, 64, 64>), grid, dim3(NT), 0, s, g);
  } else {
    dim3 grid((g.N + 31) / 32, (g.M + 31) / 32);
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI, 32, 32>), grid, dim3(NT), 0, s, g);
  }
  return hipGetLastError();
}

// Split-K plain FP32 product into C [M][N] (ldc = N): when the 32x32 grid
// has fewer than four workgroups per CU, K is split in up to SPLIT_MAX
// slices of at least 512 (multiples of 64), each slice's product goes to the
// workspace, and k_splitsum adds the slices in order (deterministic).
constexpr int SPLIT_MAX = 8;
template <bool TA, bool TB>
hipError_t gemm_splitk(GemmArgs g, float* ws, hipStream_t s) {
  const long long wgs = (long long)((g.N + 31) / 32) * ((g.M + 31) / 32);
  int S = (int)std::min<long long>(SPLIT_MAX, (1024 + wgs - 1) / wgs);
  S = std::min(S, std::max(1, g.K / 512));
  if (S <= 1 || wgs >= 1024 || g.ldc != g.N) return gemm<TA, TB, EPI_F32>(g, s);
  const int kc = ((g.K + S - 1) / S + 63) / 64 * 64;
  S = (g.K + kc - 1) / kc;
  float* C = g.C;
  g.C = ws;
  g.kc = kc;
  g.cz = (long long)g.M * g.N;
  dim3 grid((g.N + 31) / 32, (g.M + 31) / 32, S);
  hipLaunchKernelGGL((k_gemm<TA, TB, EPI_F32, 32, 32>), grid, dim3(NT), 0, s, g);
  const long long count = (long long)g.M * g.N;
  hipLaunchKernelGGL(k_splitsum, dim3(blocks(count, 256)), dim3(256), 0, s, ws, S, count, C);
  return hipGetLastError();
}

// ------------------------------------------------------------ inference
// k_fwd<EPI, KS>: one layer of inference, C(m, n) = sum_k A(m, k) Bt(n, k),
// both operands k-contiguous (Bt is the layer's weight matrix transposed),
// so every MFMA fragment is one 16-byte load straight from memory: no LDS
// staging and no barrier inside the K loop.  A workgroup covers 32 rows x
// 64 * (4 / KS) columns; its four waves hold 32 x 64 tiles (2 x 4 MFMA
// tiles) and split K in KS interleaved phases of 32-wide steps, two steps in
// flight per wave.  The partial tiles meet in LDS, are added in phase order
// (deterministic, and independent of M), and leave through the epilogue in
// row order (coalesced stores).
struct FwdArgs {
  const __bf16* A;   // [M][lda]
  const __bf16* Bt;  // [N][ldb]
  long long lda, ldb;
  int M, N, K;
  const float* bias;
  float* C;  // EPI_BIAS / EPI_BIAS_EXP
  long long ldc;
  __bf16* H;  // EPI_BIAS_GELU
  long long ldh;
  float scale;
};

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

template <int EPI, int KS>
__global__ void __launch_bounds__(NT) k_fwd(FwdArgs g) {
  constexpr int CW = 4 / KS, TNW = 64 * CW, LDR = 68;
  __shared__ float red[4][32 * LDR];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int cw = w % CW, kw = w / CW;
  const int m0 = blockIdx.y * 32, nb = blockIdx.x * TNW, n0 = nb + cw * 64;
  const int r16 = lane & 15, kg = (lane >> 4) * 8;
  const __bf16* ap[2];
  const __bf16* bp[4];
#pragma unroll
  for (int i = 0; i < 2; i++) ap[i] = g.A + (long long)min(m0 + i * 16 + r16, g.M - 1) * g.lda + kg;
#pragma unroll
  for (int j = 0; j < 4; j++) bp[j] = g.Bt + (long long)min(n0 + j * 16 + r16, g.N - 1) * g.ldb + kg;
  // the bias joins the first K phase's partial tile before the exchange, so
  // the store loop below holds no load (a load there would wait for the
  // stores in flight)
  float bj[4];
#pragma unroll
  for (int j = 0; j < 4; j++) bj[j] = kw == 0 ? g.bias[min(n0 + j * 16 + r16, g.N - 1)] : 0.0f;
  const int ns = (g.K + 31) / 32;
  f32x4 acc[2][4];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  u32x4 a0[2], b0[4], a1[2], b1[4];
  // Branch-free loads, their values used only by the MFMAs (a load under a
  // branch, or a select right after it, makes the compiler wait for the
  // loads in flight): rows past M read row M-1 and columns past N row N-1
  // (neither is stored); a k-chunk past K reads the row's last chunk and
  // its A fragment is zeroed at the MFMA (a zero times finite weights adds
  // nothing).
  const int klim = g.K - 8 - kg;  // last valid chunk offset of this lane (K >= 8)
  auto load = [&](u32x4* a, u32x4* b, int s) {
    const int ko = max(min(s * 32, klim), -kg);
#pragma unroll
    for (int i = 0; i < 2; i++) a[i] = *reinterpret_cast<const u32x4*>(ap[i] + ko);
#pragma unroll
    for (int j = 0; j < 4; j++) b[j] = *reinterpret_cast<const u32x4*>(bp[j] + ko);
  };
  auto mma = [&](const u32x4* a, const u32x4* b, int s) {
    const bool ok = s * 32 <= klim;
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const bf16x8 av = __builtin_bit_cast(bf16x8, ok ? a[i] : u32x4{0, 0, 0, 0});
#pragma unroll
      for (int j = 0; j < 4; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, __builtin_bit_cast(bf16x8, b[j]), acc[i][j], 0, 0, 0);
    }
  };
  load(a0, b0, kw);
  __builtin_amdgcn_sched_barrier(0);
  load(a1, b1, kw + KS);
  __builtin_amdgcn_sched_barrier(0);
  // program order kept (sched_barrier): each buffer's loads go out right
  // after its MFMAs, so the MFMAs of one buffer wait only for that buffer
  for (int s = kw; s < ns; s += 2 * KS) {
    mma(a0, b0, s);
    __builtin_amdgcn_sched_barrier(0);
    load(a0, b0, s + 2 * KS);
    __builtin_amdgcn_sched_barrier(0);
    mma(a1, b1, s + KS);
    __builtin_amdgcn_sched_barrier(0);
    load(a1, b1, s + 3 * KS);
    __builtin_amdgcn_sched_barrier(0);
  }
  // lane holds rows 4 * (lane / 16) + r of each 16x16 tile, column lane % 16
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) red[w][(i * 16 + (lane >> 4) * 4 + r) * LDR + j * 16 + r16] = acc[i][j][r] + bj[j];
  __syncthreads();
  // each thread keeps one column (NT is a multiple of TNW)
  constexpr int RS = NT / TNW;
  const int col = t % TNW, n = nb + col, cc = col % 64;
#pragma unroll
  for (int q = 0; q < 32 / RS; q++) {
    const int row = t / TNW + q * RS, m = m0 + row;
    if (m >= g.M || n >= g.N) continue;
    float v = red[col / 64][row * LDR + cc];
#pragma unroll
    for (int p = 1; p < KS; p++) v += red[col / 64 + p * CW][row * LDR + cc];
    const float z = v;
    if constexpr (EPI == EPI_BIAS_GELU) {
      g.H[(long long)m * g.ldh + n] = (__bf16)gelu(z);
    } else if constexpr (EPI == EPI_BIAS) {
      g.C[(long long)m * g.ldc + n] = z;
    } else {
      static_assert(EPI == EPI_BIAS_EXP, "inference epilogues: bias + GELU, bias, bias + exp");
      g.C[(long long)m * g.ldc + n] = expf(fminf(z, 80.0f)) * g.scale;
    }
  }
}

// k_fwdp<EPI, MI, NJ>: layers with K <= 256 (the hidden layer's width).
// Each wave keeps its 16*NJ-column slice of the transposed weights in
// registers (8 steps of 32 k x NJ column tiles of 16) for the whole kernel
// and takes 16*MI-row tiles of A in turn (blockIdx.y, blockIdx.y +
// gridDim.y, ...): the weights are read once per wave instead of once per
// row tile.  A workgroup's four waves hold four adjacent slices and read the
// same A rows.  A tile's result leaves through the wave's own LDS region in
// row order (coalesced stores), the bias added on the way.  No workgroup
// barrier: waves whose slice lies past N leave at once.  The tile shape only
// changes which wave computes a row, not its arithmetic (each 16x16 output
// tile is the same eight MFMAs in the same order).
template <int EPI, int MI, int NJ>
__global__ void __launch_bounds__(NT) k_fwdp(FwdArgs g) {
  constexpr int TR = 16 * MI, TC = 16 * NJ, LDR = TC + 4;
  constexpr int RPS = 64 / TC;  // rows per store step (one column per lane)
  __shared__ float tile[4][TR * LDR];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int n0 = (blockIdx.x * 4 + w) * TC;
  if (n0 >= g.N) return;
  const int r16 = lane & 15, kg = (lane >> 4) * 8;
  const int klim = g.K - 8 - kg;  // last valid chunk offset of this lane (K >= 8)
  u32x4 bf[8][NJ];
#pragma unroll
  for (int s = 0; s < 8; s++) {
    const int ko = max(min(s * 32, klim), -kg);
#pragma unroll
    for (int j = 0; j < NJ; j++)
      bf[s][j] = *reinterpret_cast<const u32x4*>(g.Bt + (long long)min(n0 + j * 16 + r16, g.N - 1) * g.ldb + kg + ko);
  }
  const int cl = lane % TC, nc = n0 + cl;  // this lane's output column in the epilogue
  const float bias = g.bias[min(nc, g.N - 1)];
  float* T = tile[w];
  const int nrt = (g.M + TR - 1) / TR;
  for (int rt = blockIdx.y; rt < nrt; rt += gridDim.y) {
    const int m0 = rt * TR;
    u32x4 af[8][MI];
#pragma unroll
    for (int i = 0; i < MI; i++) {
      const __bf16* ap = g.A + (long long)min(m0 + i * 16 + r16, g.M - 1) * g.lda + kg;
#pragma unroll
      for (int s = 0; s < 8; s++) af[s][i] = *reinterpret_cast<const u32x4*>(ap + max(min(s * 32, klim), -kg));
    }
    f32x4 acc[MI][NJ];
#pragma unroll
    for (int i = 0; i < MI; i++)
#pragma unroll
      for (int j = 0; j < NJ; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 8; s++) {
      const bool ok = s * 32 <= klim;  // k past K: the A fragment is zeroed
#pragma unroll
      for (int i = 0; i < MI; i++) {
        const bf16x8 av = __builtin_bit_cast(bf16x8, ok ? af[s][i] : u32x4{0, 0, 0, 0});
#pragma unroll
        for (int j = 0; j < NJ; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, __builtin_bit_cast(bf16x8, bf[s][j]), acc[i][j], 0, 0,
                                                               0);
      }
    }
    // lane holds rows 4 * (lane / 16) + r of each 16x16 tile, column lane % 16
#pragma unroll
    for (int i = 0; i < MI; i++)
#pragma unroll
      for (int j = 0; j < NJ; j++)
#pragma unroll
        for (int r = 0; r < 4; r++) T[(i * 16 + (lane >> 4) * 4 + r) * LDR + j * 16 + r16] = acc[i][j][r];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (nc < g.N) {
#pragma unroll 4
      for (int q = 0; q < TR / RPS; q++) {
        const int row = lane / TC + q * RPS, m = m0 + row;
        if (m >= g.M) break;
        const float z = T[row * LDR + cl] + bias;
        if constexpr (EPI == EPI_BIAS_GELU) {
          g.H[(long long)m * g.ldh + nc] = (__bf16)gelu(z);
        } else if constexpr (EPI == EPI_BIAS) {
          g.C[(long long)m * g.ldc + nc] = z;
        } else {
          static_assert(EPI == EPI_BIAS_EXP, "inference epilogues: bias + GELU, bias, bias + exp");
          g.C[(long long)m * g.ldc + nc] = expf(fminf(z, 80.0f)) * g.scale;
        }
      }
    }
    // the next tile's LDS writes come after every lane's reads
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

template <int EPI, int MI, int NJ>
hipError_t fwdp(const FwdArgs& g, hipStream_t s) {
  const int gx = (g.N + 64 * NJ - 1) / (64 * NJ), nrt = (g.M + 16 * MI - 1) / (16 * MI);
  const int gy = std::max(1, std::min(nrt, (512 + gx - 1) / gx));
  hipLaunchKernelGGL((k_fwdp<EPI, MI, NJ>), dim3(gx, gy), dim3(NT), 0, s, g);
  return hipGetLastError();
}

// k_fwdw<EPI>: layers with 256 < K <= 1280 (the input layer: bh * (bw/2+1)
// log-power features).  A workgroup of eight waves covers 32 rows x 64
// columns; wave p takes the 32-wide k steps p, p + 8, ..., at most five,
// and issues all of their loads at once (one memory round trip per wave
// instead of one per step).  The eight partial tiles meet in LDS and are
// added in phase order (deterministic, independent of M); the bias joins
// phase 0's tile; one column per thread in the coalesced epilogue.
constexpr int NTW = 512, WSTEPS = 5;
template <int EPI>
__global__ void __launch_bounds__(NTW) k_fwdw(FwdArgs g) {
  constexpr int LDR = 68;
  __shared__ float red[8][32 * LDR];
  const int t = threadIdx.x, lane = t & 63, p = t >> 6;
  const int m0 = blockIdx.y * 32, n0 = blockIdx.x * 64;
  const int r16 = lane & 15, kg = (lane >> 4) * 8;
  const int ns = (g.K + 31) / 32, klim = g.K - 8 - kg;
  u32x4 af[WSTEPS][2], bf[WSTEPS][4];
#pragma unroll
  for (int q = 0; q < WSTEPS; q++) {
    const int ko = max(min((p + 8 * q) * 32, klim), -kg);
#pragma unroll
    for (int i = 0; i < 2; i++)
      af[q][i] = *reinterpret_cast<const u32x4*>(g.A + (long long)min(m0 + i * 16 + r16, g.M - 1) * g.lda + kg + ko);
#pragma unroll
    for (int j = 0; j < 4; j++)
      bf[q][j] = *reinterpret_cast<const u32x4*>(g.Bt + (long long)min(n0 + j * 16 + r16, g.N - 1) * g.ldb + kg + ko);
  }
  float bj[4];
#pragma unroll
  for (int j = 0; j < 4; j++) bj[j] = p == 0 ? g.bias[min(n0 + j * 16 + r16, g.N - 1)] : 0.0f;
  f32x4 acc[2][4];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int q = 0; q < WSTEPS; q++) {
    const int s = p + 8 * q;
    const bool ok = s < ns && s * 32 <= klim;  // steps past K add zeros
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const bf16x8 av = __builtin_bit_cast(bf16x8, ok ? af[q][i] : u32x4{0, 0, 0, 0});
#pragma unroll
      for (int j = 0; j < 4; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, __builtin_bit_cast(bf16x8, bf[q][j]), acc[i][j], 0, 0,
                                                             0);
    }
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) red[p][(i * 16 + (lane >> 4) * 4 + r) * LDR + j * 16 + r16] = acc[i][j][r] + bj[j];
  __syncthreads();
  const int col = t & 63, n = n0 + col;
#pragma unroll
  for (int q = 0; q < 32 * 64 / NTW; q++) {
    const int row = (t >> 6) + q * (NTW / 64), m = m0 + row;
    if (m >= g.M || n >= g.N) continue;
    float v = red[0][row * LDR + col];
#pragma unroll
    for (int ph = 1; ph < 8; ph++) v += red[ph][row * LDR + col];
    if constexpr (EPI == EPI_BIAS_GELU) {
      g.H[(long long)m * g.ldh + n] = (__bf16)gelu(v);
    } else if constexpr (EPI == EPI_BIAS) {
      g.C[(long long)m * g.ldc + n] = v;
    } else {
      static_assert(EPI == EPI_BIAS_EXP, "inference epilogues: bias + GELU, bias, bias + exp");
      g.C[(long long)m * g.ldc + n] = expf(fminf(v, 80.0f)) * g.scale;
    }
  }
}

// K <= 256: weights held in registers (k_fwdp), at most about 512
// workgroups;
// K <= 1280: eight K phases with every load in flight at once (k_fwdw);
// otherwise (64x64 blocks: 2,112 inputs) four K phases of two steps in
// flight (k_fwd).
template <int EPI>
hipError_t fwd(const FwdArgs& g, hipStream_t s) {
  if (g.K <= 256) {
    // the largest wave tile that still gives every SIMD a wave (1,024 waves)
    auto waves = [&](int mi, int nj) { return (long long)((g.N + 16 * nj - 1) / (16 * nj)) * ((g.M + 16 * mi - 1) / (16 * mi)); };
    if (waves(2, 4) >= 1024) return fwdp<EPI, 2, 4>(g, s);
    if (waves(1, 4) >= 1024) return fwdp<EPI, 1, 4>(g, s);
    return fwdp<EPI, 1, 2>(g, s);
  }
  if (g.K <= 8 * WSTEPS * 32) {
    hipLaunchKernelGGL((k_fwdw<EPI>), dim3((g.N + 63) / 64, (g.M + 31) / 32), dim3(NTW), 0, s, g);
    return hipGetLastError();
  }
  hipLaunchKernelGGL((k_fwd<EPI, 4>), dim3((g.N + 63) / 64, (g.M + 31) / 32), dim3(NT), 0, s, g);
  return hipGetLastError();
}

// dst[n][k] = src[k][n] for the three weight matrices at once (blockIdx.z),
// through 32 x 32 LDS tiles.
struct TransposeSet {
  const __bf16* src[3];
  __bf16* dst[3];
  int K[3], N[3];
};
__global__ void __launch_bounds__(256) k_transpose(TransposeSet ts) {
  __shared__ __bf16 tile[32][34];
  const int z = blockIdx.z, K = ts.K[z], N = ts.N[z];
  const int k0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
  if (k0 >= K || n0 >= N) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8)
    if (k0 + r < K && n0 + tx < N) tile[r][tx] = ts.src[z][(long long)(k0 + r) * N + n0 + tx];
  __syncthreads();
  for (int r = ty; r < 32; r += 8)
    if (n0 + r < N && k0 + tx < K) ts.dst[z][(long long)(n0 + r) * K + k0 + tx] = tile[tx][r];
}

GemmArgs args(const __bf16* A, long long lda, const __bf16* B, long long ldb, int M, int N, int K) {
  GemmArgs g{};
  g.A = A;
  g.lda = lda;
  g.B = B;
  g.ldb = ldb;
  g.M = M;
  g.N = N;
  g.K = K;
  g.scale = 1.0f;
  g.kc = K;
  g.cz = 0;
  return g;
}

}  // namespace nf3d_mlp_k

using namespace nf3d_mlp_k;

namespace {

struct MlpFail {
  std::string msg;
};
void mcheck(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    (void)hipGetLastError();  // cleared, so later launch checks do not report it again
    throw MlpFail{std::string(what) + ": " + hipGetErrorString(e)};
  }
}
void merr(char* err, size_t len, const std::string& m) {
  if (err && len) {
    size_t n = std::min(len - 1, m.size());
    memcpy(err, m.data(), n);
    err[n] = 0;
  }
}
template <typename T>
T* dalloc(size_t count) {
  void* p = nullptr;
  mcheck(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc");
  return static_cast<T*>(p);
}

}  // namespace

struct nf3d_mlp {
  int n_in = 0, nh = 0, n_out = 0, max_batch = 0, device = 0;
  size_t o_w1 = 0, o_b1 = 0, o_w2 = 0, o_b2 = 0, o_w3 = 0, o_b3 = 0, np = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  __bf16* pb = nullptr;
  // the three weight matrices transposed ([N][K], k contiguous), BF16: the
  // inference operands, rewritten whenever pb is
  __bf16* pbt = nullptr;
  size_t t_w1 = 0, t_w2 = 0, t_w3 = 0;
  float *mu = nullptr, *isd = nullptr;
  float eps = 1e-10f;
  // training workspace, max_batch rows
  float *z1 = nullptr, *z2 = nullptr, *y = nullptr, *gf = nullptr, *part = nullptr, *lpart = nullptr;
  float* wsk = nullptr;  // split-K partials of the weight gradients
  __bf16 *h1 = nullptr, *h2 = nullptr, *d3 = nullptr, *d2 = nullptr, *d1 = nullptr;
  int step = 0;
  std::mutex mu_lock;

  ~nf3d_mlp() {
    for (void* q : {(void*)p, (void*)g, (void*)m, (void*)v, (void*)pb, (void*)mu, (void*)isd, (void*)z1, (void*)z2,
                    (void*)y, (void*)gf, (void*)part, (void*)lpart, (void*)h1, (void*)h2, (void*)d3, (void*)d2,
                    (void*)d1, (void*)wsk, (void*)pbt})
      (void)hipFree(q);
  }

  void init(int ni, int h, int no, int mb) {
    if (ni <= 0 || h <= 0 || no <= 0 || mb <= 0 || ni % 8 || h % 8 || no % 8)
      throw MlpFail{"layer widths must be positive multiples of 8"};
    n_in = ni;
    nh = h;
    n_out = no;
    max_batch = mb;
    mcheck(hipGetDevice(&device), "hipGetDevice");
    // every matrix starts at a multiple of 8 elements (16-byte BF16 loads)
    auto take = [&](size_t n) {
      size_t o = np;
      np += (n + 7) / 8 * 8;
      return o;
    };
    o_w1 = take((size_t)ni * h);
    o_b1 = take(h);
    o_w2 = take((size_t)h * h);
    o_b2 = take(h);
    o_w3 = take((size_t)h * no);
    o_b3 = take(no);
    p = dalloc<float>(np);
    g = dalloc<float>(np);
    m = dalloc<float>(np);
    v = dalloc<float>(np);
    pb = dalloc<__bf16>(np);
    t_w1 = 0;
    t_w2 = (size_t)ni * h;
    t_w3 = t_w2 + (size_t)h * h;
    pbt = dalloc<__bf16>(t_w3 + (size_t)h * no);
    mcheck(hipMemset(p, 0, np * 4), "hipMemset");
    mcheck(hipMemset(g, 0, np * 4), "hipMemset");
    mcheck(hipMemset(m, 0, np * 4), "hipMemset");
    mcheck(hipMemset(v, 0, np * 4), "hipMemset");
    mcheck(hipMemset(pb, 0, np * 2), "hipMemset");
    mcheck(hipMemset(pbt, 0, (t_w3 + (size_t)h * no) * 2), "hipMemset");
    mu = dalloc<float>(ni);
    isd = dalloc<float>(ni);
    std::vector<float> zero(ni, 0.0f), one(ni, 1.0f);
    mcheck(hipMemcpy(mu, zero.data(), ni * 4, hipMemcpyHostToDevice), "norm");
    mcheck(hipMemcpy(isd, one.data(), ni * 4, hipMemcpyHostToDevice), "norm");
    const size_t wide = (size_t)std::max(h, no);
    z1 = dalloc<float>((size_t)mb * h);
    z2 = dalloc<float>((size_t)mb * h);
    y = dalloc<float>((size_t)mb * no);
    gf = dalloc<float>((size_t)mb * wide);
    part = dalloc<float>((size_t)((mb + CS_ROWS - 1) / CS_ROWS) * wide);
    lpart = dalloc<float>(blocks((long long)mb * no, 256));
    wsk = dalloc<float>((size_t)SPLIT_MAX * std::max({(size_t)ni * h, (size_t)h * h, (size_t)h * no}));
    h1 = dalloc<__bf16>((size_t)mb * h);
    h2 = dalloc<__bf16>((size_t)mb * h);
    d3 = dalloc<__bf16>((size_t)mb * no);
    d2 = dalloc<__bf16>((size_t)mb * h);
    d1 = dalloc<__bf16>((size_t)mb * h);
  }

  // pbt from pb, on stream s
  void transpose_weights(hipStream_t s) {
    TransposeSet ts{};
    const int Ks[3] = {n_in, nh, nh}, Ns[3] = {nh, nh, n_out};
    const size_t from[3] = {o_w1, o_w2, o_w3}, to[3] = {t_w1, t_w2, t_w3};
    int gk = 0, gn = 0;
    for (int z = 0; z < 3; z++) {
      ts.src[z] = pb + from[z];
      ts.dst[z] = pbt + to[z];
      ts.K[z] = Ks[z];
      ts.N[z] = Ns[z];
      gk = std::max(gk, (Ks[z] + 31) / 32);
      gn = std::max(gn, (Ns[z] + 31) / 32);
    }
    hipLaunchKernelGGL(k_transpose, dim3(gn, gk, 3), dim3(256), 0, s, ts);
    mcheck(hipGetLastError(), "k_transpose");
  }

  void params_changed() {
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks((long long)np, 256)), dim3(256), 0, nullptr, p, pb, (long long)np);
    mcheck(hipGetLastError(), "k_to_bf16");
    transpose_weights(nullptr);
    mcheck(hipDeviceSynchronize(), "params");
  }

  void colsum(const float* X, int rows, int cols, float* out, hipStream_t s) {
    const int slices = (rows + CS_ROWS - 1) / CS_ROWS;  // rows <= max_batch: part holds them
    hipLaunchKernelGGL(k_colsum1, dim3((cols + 63) / 64, slices), dim3(256), 0, s, X, rows, cols, part);
    hipLaunchKernelGGL(k_colsum2, dim3(blocks(cols, 256)), dim3(256), 0, s, part, cols, slices, out);
    mcheck(hipGetLastError(), "k_colsum");
  }

  // Inference: log-PSD (exp_out = 0) or exp(log-PSD) * scale into out [n][n_out].
  void predict(const __bf16* X, int n, float* out, int exp_out, float scale, __bf16* w1, __bf16* w2,
               hipStream_t s) {
    FwdArgs a{X, pbt + t_w1, n_in, n_in, n, nh, n_in, p + o_b1, nullptr, 0, w1, nh, 1.0f};
    mcheck(fwd<EPI_BIAS_GELU>(a, s), "layer 1");
    a = FwdArgs{w1, pbt + t_w2, nh, nh, n, nh, nh, p + o_b2, nullptr, 0, w2, nh, 1.0f};
    mcheck(fwd<EPI_BIAS_GELU>(a, s), "layer 2");
    a = FwdArgs{w2, pbt + t_w3, nh, nh, n, n_out, nh, p + o_b3, out, n_out, nullptr, 0, scale};
    if (exp_out) mcheck(fwd<EPI_BIAS_EXP>(a, s), "layer 3");
    else mcheck(fwd<EPI_BIAS>(a, s), "layer 3");
  }

  // Forward with the activations kept, Whittle loss, gradients into g.
  void grad(const __bf16* X, const float* T, int n, hipStream_t s) {
    GemmArgs a = args(X, n_in, pb + o_w1, nh, n, nh, n_in);
    a.bias = p + o_b1;
    a.C = z1;
    a.ldc = nh;
    a.H = h1;
    a.ldh = nh;
    mcheck(gemm<false, false, EPI_BIAS_GELU>(a, s), "layer 1");
    a = args(h1, nh, pb + o_w2, nh, n, nh, nh);
    a.bias = p + o_b2;
    a.C = z2;
    a.ldc = nh;
    a.H = h2;
    a.ldh = nh;
    mcheck(gemm<false, false, EPI_BIAS_GELU>(a, s), "layer 2");
    a = args(h2, nh, pb + o_w3, n_out, n, n_out, nh);
    a.bias = p + o_b3;
    a.C = y;
    a.ldc = n_out;
    mcheck(gemm<false, false, EPI_BIAS>(a, s), "layer 3");
    const long long tot = (long long)n * n_out;
    hipLaunchKernelGGL(k_whittle, dim3(blocks(tot, 256)), dim3(256), 0, s, y, T, d3, gf, tot,
                       1.0f / (float)tot, lpart);
    mcheck(hipGetLastError(), "k_whittle");
    // layer 3: dW3 = H2^T dY, db3 = colsum dY, dZ2 = (dY W3^T) * gelu'(Z2)
    a = args(h2, nh, d3, n_out, nh, n_out, n);
    a.C = g + o_w3;
    a.ldc = n_out;
    mcheck(gemm_splitk<true, false>(a, wsk, s), "dW3");
    colsum(gf, n, n_out, g + o_b3, s);
    a = args(d3, n_out, pb + o_w3, n_out, n, nh, n_out);
    a.H = d2;
    a.ldh = nh;
    a.C = gf;
    a.ldc = nh;
    a.Zp = z2;
    a.ldz = nh;
    mcheck(gemm<false, true, EPI_GELU_BWD>(a, s), "dZ2");
    // layer 2
    a = args(h1, nh, d2, nh, nh, nh, n);
    a.C = g + o_w2;
    a.ldc = nh;
    mcheck(gemm_splitk<true, false>(a, wsk, s), "dW2");
    colsum(gf, n, nh, g + o_b2, s);
    a = args(d2, nh, pb + o_w2, nh, n, nh, nh);
    a.H = d1;
    a.ldh = nh;
    a.C = gf;
    a.ldc = nh;
    a.Zp = z1;
    a.ldz = nh;
    mcheck(gemm<false, true, EPI_GELU_BWD>(a, s), "dZ1");
    // layer 1
    a = args(X, n_in, d1, nh, n_in, nh, n);
    a.C = g + o_w1;
    a.ldc = nh;
    mcheck(gemm_splitk<true, false>(a, wsk, s), "dW1");
    colsum(gf, n, nh, g + o_b1, s);
  }

  double read_loss(int n, hipStream_t s) {
    const long long tot = (long long)n * n_out;
    std::vector<float> h(blocks(tot, 256));
    mcheck(hipMemcpyAsync(h.data(), lpart, h.size() * 4, hipMemcpyDeviceToHost, s), "loss");
    mcheck(hipStreamSynchronize(s), "loss");
    double sum = 0;
    for (float x : h) sum += x;
    return sum / (double)tot;
  }

  void adam(float lr, float b1, float b2, float eps_, hipStream_t s) {
    step++;
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, step)));
    hipLaunchKernelGGL(k_adam, dim3(blocks((long long)np, 256)), dim3(256), 0, s, p, g, m, v, pb, (long long)np, lr,
                       b1, b2, eps_, c1, c2);
    mcheck(hipGetLastError(), "k_adam");
    transpose_weights(s);
  }
};

namespace {

template <class F>
int guarded(char* err, size_t errlen, F&& f) {
  try {
    f();
  } catch (const MlpFail& e) {
    merr(err, errlen, e.msg);
    return -1;
  } catch (const std::bad_alloc&) {
    merr(err, errlen, "out of host memory");
    return -1;
  }
  return 0;
}

hipStream_t st(void* s) { return static_cast<hipStream_t>(s); }

}  // namespace

// Features of nseg spectra arrays (rows[s] blocks each) into consecutive
// rows of x_out, one launch.
int nf3d_mlp_features_batch(nf3d_mlp* m, const void* const* specs, const int* rows, int nseg, void* x_out,
                            void* stream) {
  if (!m || nseg < 1 || nseg > 8) return -1;
  const int d = m->n_in;
  if (d % 4 || d / 4 > 1024) return -1;
  FeatSegs sg{};
  long long total = 0;
  for (int s = 0; s < nseg; s++) {
    if (rows[s] < 0 || (rows[s] > 0 && !specs[s])) return -1;
    sg.spec[s] = static_cast<const float2*>(specs[s]);
    sg.row0[s] = (int)total;
    total += rows[s];
  }
  if (total > 0x7fffffff) return -1;
  sg.row0[nseg] = (int)total;
  sg.n = nseg;
  if (!total) return 0;
  if (!x_out) return -1;
  const int threads = (d / 4 + 63) / 64 * 64;
  hipLaunchKernelGGL(k_features_rows, dim3((unsigned)total), dim3(threads), 0, st(stream), sg,
                     static_cast<__bf16*>(x_out), d, m->mu, m->isd, m->eps);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------- C ABI
extern "C" {

int nf3d_mlp_create(int n_in, int n_hidden, int n_out, int max_batch, nf3d_mlp** out, char* err, size_t errlen) {
  if (!out) return -1;
  auto* m = new nf3d_mlp();
  int rc = guarded(err, errlen, [&] { m->init(n_in, n_hidden, n_out, max_batch); });
  if (rc) {
    delete m;
    return rc;
  }
  *out = m;
  return 0;
}

void nf3d_mlp_destroy(nf3d_mlp* m) { delete m; }

long long nf3d_mlp_param_count(const nf3d_mlp* m) {
  return (long long)m->n_in * m->nh + m->nh + (long long)m->nh * m->nh + m->nh + (long long)m->nh * m->n_out +
         m->n_out;
}

void nf3d_mlp_dims(const nf3d_mlp* m, int* o) {
  o[0] = m->n_in;
  o[1] = m->nh;
  o[2] = m->n_out;
}

// Host parameters in the order W1 [n_in][h], b1, W2 [h][h], b2, W3 [h][n_out], b3.
static void pieces(const nf3d_mlp* m, size_t off[6], size_t len[6]) {
  off[0] = m->o_w1, len[0] = (size_t)m->n_in * m->nh;
  off[1] = m->o_b1, len[1] = m->nh;
  off[2] = m->o_w2, len[2] = (size_t)m->nh * m->nh;
  off[3] = m->o_b2, len[3] = m->nh;
  off[4] = m->o_w3, len[4] = (size_t)m->nh * m->n_out;
  off[5] = m->o_b3, len[5] = m->n_out;
}

int nf3d_mlp_set_params(nf3d_mlp* m, const float* host, char* err, size_t errlen) {
  std::lock_guard<std::mutex> lock(m->mu_lock);
  return guarded(err, errlen, [&] {
    mcheck(hipSetDevice(m->device), "hipSetDevice");
    size_t off[6], len[6], at = 0;
    pieces(m, off, len);
    for (int i = 0; i < 6; i++) {
      mcheck(hipMemcpy(m->p + off[i], host + at, len[i] * 4, hipMemcpyHostToDevice), "set_params");
      at += len[i];
    }
    mcheck(hipMemset(m->m, 0, m->np * 4), "hipMemset");
    mcheck(hipMemset(m->v, 0, m->np * 4), "hipMemset");
    m->step = 0;
    m->params_changed();
  });
}

static int copy_out(nf3d_mlp* m, const float* dev, float* host, char* err, size_t errlen) {
  std::lock_guard<std::mutex> lock(m->mu_lock);
  return guarded(err, errlen, [&] {
    mcheck(hipSetDevice(m->device), "hipSetDevice");
    mcheck(hipDeviceSynchronize(), "sync");
    size_t off[6], len[6], at = 0;
    pieces(m, off, len);
    for (int i = 0; i < 6; i++) {
      mcheck(hipMemcpy(host + at, dev + off[i], len[i] * 4, hipMemcpyDeviceToHost), "copy");
      at += len[i];
    }
  });
}

int nf3d_mlp_get_params(nf3d_mlp* m, float* host, char* err, size_t errlen) {
  return copy_out(m, m->p, host, err, errlen);
}

int nf3d_mlp_get_grads(nf3d_mlp* m, float* host, char* err, size_t errlen) {
  return copy_out(m, m->g, host, err, errlen);
}

int nf3d_mlp_set_input_norm(nf3d_mlp* m, const float* mu, const float* isd, float eps, char* err, size_t errlen) {
  std::lock_guard<std::mutex> lock(m->mu_lock);
  return guarded(err, errlen, [&] {
    mcheck(hipSetDevice(m->device), "hipSetDevice");
    mcheck(hipMemcpy(m->mu, mu, m->n_in * 4, hipMemcpyHostToDevice), "norm");
    mcheck(hipMemcpy(m->isd, isd, m->n_in * 4, hipMemcpyHostToDevice), "norm");
    m->eps = eps;
  });
}

int nf3d_mlp_features(nf3d_mlp* m, const void* spec, int n, void* x_out, void* stream) {
  return nf3d_mlp_features_batch(m, &spec, &n, 1, x_out, stream);
}

int nf3d_to_bf16(const float* x, void* out, long long count, void* stream) {
  if (count < 0) return -1;
  if (!count) return 0;
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(count, 256)), dim3(256), 0, st(stream), x, static_cast<__bf16*>(out),
                     count);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int nf3d_gather_rows(const void* src, const int* idx, void* dst, int rows, int row_bytes, void* stream) {
  if (rows < 0 || row_bytes <= 0 || row_bytes % 16) return -1;
  if (!rows) return 0;
  const int words = row_bytes / 16;
  hipLaunchKernelGGL(k_gather_rows, dim3(blocks((long long)rows * words, 256)), dim3(256), 0, st(stream),
                     static_cast<const uint4*>(src), idx, static_cast<uint4*>(dst), rows, words);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int nf3d_periodogram(const void* spec, float* out, long long count, void* stream) {
  if (count < 0) return -1;
  if (!count) return 0;
  hipLaunchKernelGGL(k_periodogram, dim3(blocks(count, 256)), dim3(256), 0, st(stream),
                     static_cast<const float2*>(spec), out, count);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

long long nf3d_mlp_work_bytes(const nf3d_mlp* m, int n) { return 2LL * n * m->nh * 2; }

int nf3d_mlp_predict(nf3d_mlp* m, const void* x, int n, float* out, int exp_out, float scale, void* work,
                     void* stream, char* err, size_t errlen) {
  return guarded(err, errlen, [&] {
    if (n < 0 || (n > 0 && (!x || !out || !work))) throw MlpFail{"bad arguments"};
    if (!n) return;
    __bf16* w1 = static_cast<__bf16*>(work);
    m->predict(static_cast<const __bf16*>(x), n, out, exp_out, scale, w1, w1 + (size_t)n * m->nh, st(stream));
  });
}

}  // extern "C"

// ------------------------------------------------- inference without allocation
// The hidden activations of nf3d_mlp_predict_noalloc live in this array
// instead of a caller's workspace: 2 layers x up to 16384 rows x 256 hidden
// units, BF16 (16 MiB), one copy per GPU, held until the process ends.  Host
// code never passes `scratch` itself to a kernel (on the host the name is a
// stand-in, not a device address): it looks up each device's address once
// with hipGetSymbolAddress and passes that.
namespace {
constexpr long long kScratchElems = 2LL * 16384 * 256;
constexpr int kMaxDevices = 64;
}  // namespace
static __device__ __align__(256) __bf16 scratch[kScratchElems];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `dev`'s copy of the scratch array, looked up on the
// first call on that device with the device current.
void* scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw MlpFail{"device index outside 0..63"};
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    mcheck(hipGetDevice(&cur), "hipGetDevice");
    mcheck(hipSetDevice(dev), "hipSetDevice");
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    (void)hipSetDevice(cur);
    if (e != hipSuccess || !ptr) {
      (void)hipGetLastError();
      throw MlpFail{std::string("hipGetSymbolAddress of the scratch array: ") + hipGetErrorString(e)};
    }
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int nf3d_mlp_predict_noalloc(nf3d_mlp* m, const void* x, int n, float* out, int exp_out, float scale, void* stream,
                             char* err, size_t errlen) {
  return guarded(err, errlen, [&] {
    if (!m || n < 0 || (n > 0 && (!x || !out))) throw MlpFail{"bad arguments"};
    if (2LL * n * m->nh > kScratchElems)
      throw MlpFail{"no-allocation inference holds at most 16384 rows of 256 hidden units (n * n_hidden <= " +
                    std::to_string(kScratchElems / 2) + ")"};
    if (!n) return;
    __bf16* w1 = static_cast<__bf16*>(scratch_on(m->device));
    m->predict(static_cast<const __bf16*>(x), n, out, exp_out, scale, w1, w1 + (size_t)n * m->nh, st(stream));
  });
}

int nf3d_mlp_grad(nf3d_mlp* m, const void* x, const float* t, int n, double* loss, void* stream, char* err,
                  size_t errlen) {
  std::lock_guard<std::mutex> lock(m->mu_lock);
  return guarded(err, errlen, [&] {
    if (n <= 0 || n > m->max_batch || !x || !t) throw MlpFail{"batch size outside 1..max_batch"};
    m->grad(static_cast<const __bf16*>(x), t, n, st(stream));
    if (loss) *loss = m->read_loss(n, st(stream));
  });
}

int nf3d_mlp_adam(nf3d_mlp* m, float lr, float beta1, float beta2, float eps, void* stream, char* err,
                  size_t errlen) {
  std::lock_guard<std::mutex> lock(m->mu_lock);
  return guarded(err, errlen, [&] { m->adam(lr, beta1, beta2, eps, st(stream)); });
}

int nf3d_mlp_train_step(nf3d_mlp* m, const void* x, const float* t, int n, float lr, double* loss, void* stream,
                        char* err, size_t errlen) {
  std::lock_guard<std::mutex> lock(m->mu_lock);
  return guarded(err, errlen, [&] {
    if (n <= 0 || n > m->max_batch || !x || !t) throw MlpFail{"batch size outside 1..max_batch"};
    m->grad(static_cast<const __bf16*>(x), t, n, st(stream));
    if (loss) *loss = m->read_loss(n, st(stream));
    m->adam(lr, 0.9f, 0.999f, 1e-8f, st(stream));
  });
}

}  // extern "C"
