// Device code of the per-block kernel k_block<BW,BH,NT,MODE,BT> (gather,
// register r2c, spectral pass, c2r), shared by the per-shape translation
// units (nf3d_shape.hip is compiled once per block shape).
#pragma once
#include "nf3d_fft.h"
#include "nf3d_internal.h"

namespace nf3d {

// ----------------------------------------------------------------- helpers
__device__ __forceinline__ int reflect101(int v, int n) {
  v = v < 0 ? -v : v;
  v = v >= n ? 2 * n - 2 - v : v;
  return v;
}

__device__ __forceinline__ float load_px(const uint8_t* row, int x, int pix) {
  if (pix == 1) return (float)row[x];
  if (pix == 2) return (float)reinterpret_cast<const uint16_t*>(row)[x];
  return reinterpret_cast<const float*>(row)[x];
}

// psd-based sharpen and dehalo factors (src/code_impl/Sharpen.cpp:102-105):
// the reference evaluates the square root in double.
__device__ __forceinline__ float sharpen_factor(const BlockJob& J, float psd, int k) {
#pragma clang fp contract(off)
  float a = J.sharpen * J.wsharpen[k];
  float q = psd * J.smax_n / ((psd + J.smin_n) * (psd + J.smax_n));
  return (float)(1.0 + (double)a * sqrt((double)q));
}
__device__ __forceinline__ float dehalo_factor(const BlockJob& J, float psd, int k) {
#pragma clang fp contract(off)
  float s = psd + J.ht2n;
  return s / (s + J.dehalo * J.wdehalo[k] * psd);
}

// Limited Wiener gain applied in place (code_impl_C.h:111-117).
__device__ __forceinline__ void wiener3d(float& re, float& im, float sig, float lowlimit) {
#pragma clang fp contract(off)
  float psd = re * re + im * im + 1.0e-15f;
  float f = (psd - sig) / psd;
  f = f > lowlimit ? f : lowlimit;
  re *= f;
  im *= f;
}

// ---------------------------------------------------------------- spectral
// Column of the current block: col[i] = X[i][c], i < BH, c = lane's column.

template <int BH, int COLS>
__device__ void apply2d(float2 (&col)[BH], const BlockJob& J, int c, float dc) {
#pragma clang fp contract(off)
  float gf = J.use_degrid ? J.degrid * dc / J.grid[0].x : 0.0f;
  static_for<0, BH>([&](auto ii) {
    constexpr int i = decltype(ii)::value;
    int k = i * COLS + c;
    float g0 = 0.0f, g1 = 0.0f;
    if (J.use_degrid) {
      float2 g = J.grid[k];
      g0 = gf * g.x;
      g1 = gf * g.y;
    }
    float cr = col[i].x - g0, ci = col[i].y - g1;
    float psd = cr * cr + ci * ci + 1e-15f;
    float sig = J.use_pattern ? J.pattern2d[k] : J.sigma2;
    float f = (psd - sig) / psd;
    f = f > J.lowlimit ? f : J.lowlimit;
    if (!J.use_pattern) {  // Apply.cpp:33-46 variant selection (Apply.cpp:57-64)
      if (J.dehalo == 0.0f) f *= sharpen_factor(J, psd, k);
      else if (J.sharpen == 0.0f) f *= dehalo_factor(J, psd, k);
      else f *= sharpen_factor(J, psd, k) * dehalo_factor(J, psd, k);
    }
    col[i].x = cr * f + g0;
    col[i].y = ci * f + g1;
  });
}

template <int BH, int COLS>
__device__ void sharpen_pass(float2 (&col)[BH], const BlockJob& J, int c, float dc) {
#pragma clang fp contract(off)
  float gf = J.use_degrid ? J.degrid * dc / J.grid[0].x : 0.0f;
  static_for<0, BH>([&](auto ii) {
    constexpr int i = decltype(ii)::value;
    int k = i * COLS + c;
    float g0 = 0.0f, g1 = 0.0f;
    if (J.use_degrid) {
      float2 g = J.grid[k];
      g0 = gf * g.x;
      g1 = gf * g.y;
    }
    float cr = col[i].x - g0, ci = col[i].y - g1;
    float psd = cr * cr + ci * ci + 1.0e-15f;
    float f;
    if (J.dehalo == 0.0f) f = sharpen_factor(J, psd, k);
    else if (J.sharpen == 0.0f) f = dehalo_factor(J, psd, k);
    else f = sharpen_factor(J, psd, k) * dehalo_factor(J, psd, k);
    col[i].x = cr * f + g0;
    col[i].y = ci * f + g1;
  });
}

// Temporal DFT of length BT over frames spec_in[0..BT-1] (oldest first),
// Wiener per temporal bin, degrid x BT on the DC bin, inverse at the current
// frame (src/code_impl/Apply.cpp:67-235, same float evaluation order).
template <int BT, int BH, int COLS>
__device__ void apply3d(float2 (&col)[BH], const BlockJob& J, size_t blk_off, int c) {
#pragma clang fp contract(off)
  const float2* in[5] = {J.spec_in[0], J.spec_in[1], J.spec_in[2], J.spec_in[3], J.spec_in[4]};
  float dc = J.spec_cur[blk_off].x;
  float gf = J.use_degrid ? J.degrid * dc / J.grid[0].x : 0.0f;
  const float ll = J.lowlimit;
  static_for<0, BH>([&](auto ii) {
    constexpr int i = decltype(ii)::value;
    int k = i * COLS + c;
    size_t o = blk_off + (size_t)k;
    float sig = J.use_pattern ? J.pattern3d[k] : J.sigma2;
    float g0 = 0.0f, g1 = 0.0f;
    if (J.use_degrid) {
      float2 g = J.grid[k];
      g0 = gf * g.x * (float)BT;
      g1 = gf * g.y * (float)BT;
    }
    float2 r;
    if constexpr (BT == 2) {
      float2 p = in[0][o], q = in[1][o];  // prev, cur
      float f0r = q.x + p.x - g0, f0i = q.y + p.y - g1;
      float f1r = q.x - p.x, f1i = q.y - p.y;
      wiener3d(f0r, f0i, sig, ll);
      wiener3d(f1r, f1i, sig, ll);
      r.x = (f0r + f1r + g0) * 0.5f;
      r.y = (f0i + f1i + g1) * 0.5f;
    } else if constexpr (BT == 3) {
      constexpr float s120 = 0.86602540378443864676372317075294f;
      constexpr float third = 1.0f / 3.0f;
      float2 p = in[0][o], q = in[1][o], n = in[2][o];  // prev, cur, next
      float pnr = p.x + n.x, pni = p.y + n.y;
      float fcr = q.x + pnr - g0, fci = q.y + pni - g1;
      float di = s120 * (p.y - n.y);
      float dr = s120 * (n.x - p.x);
      float fpr = q.x - 0.5f * pnr + di;
      float fnr = q.x - 0.5f * pnr - di;
      float fpi = q.y - 0.5f * pni + dr;
      float fni = q.y - 0.5f * pni - dr;
      wiener3d(fcr, fci, sig, ll);
      wiener3d(fpr, fpi, sig, ll);
      wiener3d(fnr, fni, sig, ll);
      r.x = (fcr + fpr + fnr + g0) * third;
      r.y = (fci + fpi + fni + g1) * third;
    } else if constexpr (BT == 4) {
      float2 p2 = in[0][o], p = in[1][o], q = in[2][o], n = in[3][o];
      float fp2r = (q.x + p2.x) - (p.x + n.x);
      float fp2i = (q.y + p2.y) - (p.y + n.y);
      float fpr = (q.x - p2.x) + (p.y - n.y);
      float fpi = (q.y - p2.y) - (p.x - n.x);
      float fcr = (q.x + p2.x) + (p.x + n.x) - g0;
      float fci = (q.y + p2.y) + (p.y + n.y) - g1;
      float fnr = (q.x - p2.x) - (p.y - n.y);
      float fni = (q.y - p2.y) + (p.x - n.x);
      wiener3d(fp2r, fp2i, sig, ll);
      wiener3d(fpr, fpi, sig, ll);
      wiener3d(fcr, fci, sig, ll);
      wiener3d(fnr, fni, sig, ll);
      r.x = ((fp2r + fpr) + (fcr + fnr) + g0) * 0.25f;
      r.y = ((fp2i + fpi) + (fci + fni) + g1) * 0.25f;
    } else {
      constexpr float s72 = 0.95105651629515357211643933337938f;
      constexpr float c72 = 0.30901699437494742410229341718282f;
      constexpr float s144 = 0.58778525229247312916870595463907f;
      constexpr float c144 = -0.80901699437494742410229341718282f;
      float2 p2 = in[0][o], p = in[1][o], q = in[2][o], n = in[3][o], n2 = in[4][o];
      float sum = (p2.x + n2.x) * c72 + (p.x + n.x) * c144 + q.x;
      float dif = (-p2.y + n2.y) * s72 + (p.y - n.y) * s144;
      float fp2r = sum + dif, fn2r = sum - dif;
      sum = (p2.y + n2.y) * c72 + (p.y + n.y) * c144 + q.y;
      dif = (p2.x - n2.x) * s72 + (-p.x + n.x) * s144;
      float fp2i = sum + dif, fn2i = sum - dif;
      sum = (p2.x + n2.x) * c144 + (p.x + n.x) * c72 + q.x;
      dif = (p2.y - n2.y) * s144 + (p.y - n.y) * s72;
      float fpr = sum + dif, fnr = sum - dif;
      sum = (p2.y + n2.y) * c144 + (p.y + n.y) * c72 + q.y;
      dif = (-p2.x + n2.x) * s144 + (-p.x + n.x) * s72;
      float fpi = sum + dif, fni = sum - dif;
      float fcr = (p2.x + p.x) + q.x + (n.x + n2.x) - g0;
      float fci = (p2.y + p.y) + q.y + (n.y + n2.y) - g1;
      wiener3d(fp2r, fp2i, sig, ll);
      wiener3d(fpr, fpi, sig, ll);
      wiener3d(fcr, fci, sig, ll);
      wiener3d(fnr, fni, sig, ll);
      wiener3d(fn2r, fn2i, sig, ll);
      r.x = ((fp2r + fpr) + (fcr + fnr) + fn2r + g0) * 0.2f;
      r.y = ((fp2i + fpi) + (fci + fni) + fn2i + g1) * 0.2f;
    }
    col[i] = r;
  });
}

// Kalman recursion per re/im (src/code_impl/Kalman.cpp:27-64); the state
// lives in HBM in the spectrum layout and is updated in place.
template <int BH, int COLS>
__device__ void kalman(float2 (&col)[BH], const BlockJob& J, size_t blk_off, int c) {
#pragma clang fp contract(off)
  static_for<0, BH>([&](auto ii) {
    constexpr int i = decltype(ii)::value;
    int k = i * COLS + c;
    size_t o = blk_off + (size_t)k;
    float sigma;
    if (J.use_pattern) {
      sigma = J.pattern2d[k];
      if (sigma < 1e-15f) sigma = 1e-15f;
    } else {
      sigma = J.sigma2d;
    }
    float2 cur = col[i];
    float2 prev = J.kal_last[o];
    float2 cv = J.kal_covar[o], cp = J.kal_cproc[o];
    float2 out;
    float dr = cur.x - prev.x, di = cur.y - prev.y;
    float thr = sigma * J.kratio2;
    if (dr * dr > thr || di * di > thr) {
      cv = make_float2(sigma, sigma);
      cp = make_float2(sigma, sigma);
      out = cur;
    } else {
      float sr = cv.x + cp.x, si = cv.y + cp.y;
      float gr = sr / (sr + sigma), gi = si / (si + sigma);
      cp.x = gr * gr * sigma;
      cp.y = gi * gi * sigma;
      cv.x = (1 - gr) * sr;
      cv.y = (1 - gi) * si;
      out.x = gr * cur.x + (1 - gr) * prev.x;
      out.y = gi * cur.y + (1 - gi) * prev.y;
    }
    J.kal_last[o] = out;
    J.kal_covar[o] = cv;
    J.kal_cproc[o] = cp;
    col[i] = out;
  });
}

// ------------------------------------------------------------- block kernel
template <int BW, int BH, int NT>
struct Shape {
  static constexpr int PAIRS = BH / 2;
  static constexpr int COLS = BW / 2 + 1;
  static constexpr int TPB = COLS > PAIRS ? COLS : PAIRS;
  static constexpr int BPW = NT / TPB;
  static constexpr int SA = BW + 1;  // row-pair stride (float2) in the gather stage
  static constexpr int SB = BH + 2;  // column stride (float2) in the spectral stage
  static constexpr int LA = PAIRS * SA;
  static constexpr int LB = COLS * SB;
  static constexpr int LBLK = LA > LB ? LA : LB;
};

template <int BW, int BH, int NT, int MODE, int BT>
__global__ void __launch_bounds__(NT) k_block(BlockJob J) {
  using S = Shape<BW, BH, NT>;
  constexpr int COLS = S::COLS, PAIRS = S::PAIRS, BPW = S::BPW, SA = S::SA, SB = S::SB, LBLK = S::LBLK;
  __shared__ __align__(16) float2 lds[BPW * LBLK];
  __shared__ float dcs[BPW];
  __shared__ float winx[BW], winy[BH];

  const int t = threadIdx.x;
  const int blk0 = blockIdx.x * BPW;
  const int cb = t / COLS, cc = t - (t / COLS) * COLS;
  const bool col_on = cb < BPW && blk0 + cb < J.nb;
  const int rb = t / PAIRS, rp = t - (t / PAIRS) * PAIRS;
  const bool row_on = rb < BPW && blk0 + rb < J.nb;
  const size_t spec_blk = (size_t)(blk0 + cb) * BH * COLS;

  float2 col[BH];

  if constexpr (MODE != MODE_W3D) {
    for (int i = t; i < BW; i += NT) winx[i] = J.wanx[i];
    for (int i = t; i < BH; i += NT) winy[i] = J.wany[i];
    __syncthreads();
    // 1. gather
    const int stepx = BW - J.ow, stepy = BH - J.oh;
    for (int idx = t; idx < BPW * BH * BW; idx += NT) {
      int b = idx / (BH * BW);
      int rem = idx - b * (BH * BW);
      int i = rem / BW, j = rem - (rem / BW) * BW;
      int gb = blk0 + b;
      if (gb >= J.nb) continue;
      int by = gb / J.nox, bx = gb - (gb / J.nox) * J.nox;
      int ry = reflect101(by * stepy + i - J.mirh, J.Hc);
      if (J.interlaced) {
        int half = J.Hc >> 1;
        ry = ry < half ? 2 * ry : J.Hc - 1 - 2 * (ry - half);
      }
      int sx = reflect101(bx * stepx + j - J.mirw, J.Wc);
      const uint8_t* row = J.src + (long long)(ry + J.tc) * J.src_pitch;
      float p = load_px(row, sx + J.lc, J.pix) - J.base;
      float v;
      {
#pragma clang fp contract(off)
        bool xw = j < J.ow || j >= BW - J.ow;
        bool yw = i < J.oh || i >= BH - J.oh;
        float wx = winx[j], wy = winy[i];
        if (xw && yw) {
          bool yedge = (by == 0 && i < J.oh) || (by == J.noy - 1 && i >= BH - J.oh);
          bool xedge = (bx == 0 && j < J.ow) || (bx == J.nox - 1 && j >= BW - J.ow);
          if (yedge) v = xedge ? (wx * wy) * p : (wy * p) * wx;
          else v = (wx * p) * wy;
        } else if (xw) {
          v = wx * p;
        } else if (yw) {
          v = wy * p;
        } else {
          v = p;
        }
      }
      reinterpret_cast<float*>(&lds[b * LBLK + (i >> 1) * SA + j])[i & 1] = v;
    }
    __syncthreads();
    // 2. row r2c, two real rows per lane
    float2 z[BW];
    if (row_on) {
      static_for<0, BW>([&](auto j) { z[j] = lds[rb * LBLK + rp * SA + j]; });
    }
    __syncthreads();  // stage A and stage B share the block's LDS slot
    if (row_on) {
      fft<BW, -1>(z);
      static_for<0, COLS>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        constexpr int km = (BW - k) % BW;
        float2 a = z[k], b = z[km];
        float2 xa = make_float2((a.x + b.x) * 0.5f, (a.y - b.y) * 0.5f);
        float2 xb = make_float2((a.y + b.y) * 0.5f, (b.x - a.x) * 0.5f);
        float4* dst = reinterpret_cast<float4*>(&lds[rb * LBLK + k * SB + 2 * rp]);
        *dst = make_float4(xa.x, xa.y, xb.x, xb.y);
      });
    }
    __syncthreads();
    // 3. column FFT
    if (col_on) {
      static_for<0, BH>([&](auto i) { col[i] = lds[cb * LBLK + cc * SB + i]; });
      fft<BH, -1>(col);
    }
  }

  if constexpr (MODE == MODE_FWD) {
    if (col_on) {
      static_for<0, BH>([&](auto i) { J.spec_out[spec_blk + decltype(i)::value * COLS + cc] = col[i]; });
    }
    return;
  } else {
    // 4. spectral pass
    if constexpr (MODE == MODE_W2D) {
      if (col_on && cc == 0) dcs[cb] = col[0].x;
      __syncthreads();
      if (col_on) apply2d<BH, COLS>(col, J, cc, dcs[cb]);
      if (J.use_pattern) {  // pattern variant: separate sharpen pass (fft3d_engine.h:620-621)
        __syncthreads();
        if (col_on && cc == 0) dcs[cb] = col[0].x;
        __syncthreads();
        if (col_on && (J.sharpen != 0.0f || J.dehalo != 0.0f)) sharpen_pass<BH, COLS>(col, J, cc, dcs[cb]);
      }
    } else {
      if constexpr (MODE == MODE_W3D) {
        if (col_on) apply3d<BT, BH, COLS>(col, J, spec_blk, cc);
      } else if constexpr (MODE == MODE_KALMAN) {
        if (col_on) kalman<BH, COLS>(col, J, spec_blk, cc);
      }
      if (J.sharpen != 0.0f || J.dehalo != 0.0f) {
        __syncthreads();
        if (col_on && cc == 0) dcs[cb] = col[0].x;
        __syncthreads();
        if (col_on) sharpen_pass<BH, COLS>(col, J, cc, dcs[cb]);
      }
    }
    // 5. inverse: columns, then rows (c2r)
    if (col_on) fft<BH, +1>(col);
    __syncthreads();
    if (col_on) {
      static_for<0, BH>([&](auto i) { lds[cb * LBLK + cc * SB + decltype(i)::value] = col[i]; });
    }
    __syncthreads();
    if (row_on) {
      float2 z[BW];
      static_for<0, COLS>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        float4 v = *reinterpret_cast<const float4*>(&lds[rb * LBLK + k * SB + 2 * rp]);
        float ar = v.x, ai = v.y, br = v.z, bi = v.w;
        if constexpr (k == 0 || 2 * k == BW) { ai = 0.0f; bi = 0.0f; }
        z[k] = make_float2(ar - bi, ai + br);
        if constexpr (k != 0 && 2 * k != BW) z[BW - k] = make_float2(ar + bi, br - ai);
      });
      fft<BW, +1>(z);
      float* out = J.blocks + ((size_t)(blk0 + rb) * BH + 2 * rp) * BW;
      static_for<0, BW>([&](auto j) {
        out[decltype(j)::value] = z[j].x;
        out[BW + decltype(j)::value] = z[j].y;
      });
    }
  }
}

// ---------------------------------------------------------------- dispatch
template <int BW, int BH>
hipError_t launch_shape(int mode, int bt, const BlockJob& J, hipStream_t s) {
  constexpr int NT = 128;
  using S = Shape<BW, BH, NT>;
  dim3 grid((J.nb + S::BPW - 1) / S::BPW);
  switch (mode) {
    case MODE_FWD: hipLaunchKernelGGL((k_block<BW, BH, NT, MODE_FWD, 0>), grid, dim3(NT), 0, s, J); break;
    case MODE_W2D: hipLaunchKernelGGL((k_block<BW, BH, NT, MODE_W2D, 0>), grid, dim3(NT), 0, s, J); break;
    case MODE_KALMAN: hipLaunchKernelGGL((k_block<BW, BH, NT, MODE_KALMAN, 0>), grid, dim3(NT), 0, s, J); break;
    case MODE_SHARPEN: hipLaunchKernelGGL((k_block<BW, BH, NT, MODE_SHARPEN, 0>), grid, dim3(NT), 0, s, J); break;
    case MODE_W3D:
      switch (bt) {
        case 2: hipLaunchKernelGGL((k_block<BW, BH, NT, MODE_W3D, 2>), grid, dim3(NT), 0, s, J); break;
        case 3: hipLaunchKernelGGL((k_block<BW, BH, NT, MODE_W3D, 3>), grid, dim3(NT), 0, s, J); break;
        case 4: hipLaunchKernelGGL((k_block<BW, BH, NT, MODE_W3D, 4>), grid, dim3(NT), 0, s, J); break;
        case 5: hipLaunchKernelGGL((k_block<BW, BH, NT, MODE_W3D, 5>), grid, dim3(NT), 0, s, J); break;
        default: return hipErrorInvalidValue;
      }
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace nf3d
