// Fused per-tile kernel of the denoiser: k_tile<BW,BH,NT,MODE,BT,ONCHIP>.
//
// A workgroup owns a tile of up to BPW horizontally adjacent blocks of one
// block row of one plane (several planes share a launch).  Phases, with LDS
// reused between them:
//   1. gather: the tile's cover region (BH rows x (nblk-1)*stepx+BW
//      columns) is read from the plane in HBM with reflect-101 addressing,
//      16 rows per batch so every lane keeps 16-32 loads in flight;
//   2. row r2c: one lane per block row pair applies the analysis window (the
//      reference's per-region product order, src/buffer.cpp:137-378) and runs
//      a register FFT of length BW (two real rows packed as one complex row);
//   3. column FFT: one lane per spectral column, register FFT of length BH;
//      the spectrum is stored to the HBM ring when it is a frame's first
//      transform (MODE_FWD, or the newest frame of a Wiener-3D window);
//   4. spectral pass, element-wise over the tile (coalesced ring reads):
//      Wiener 2D / 3D (bt 2..5), Kalman, sharpen/dehalo, degrid, pattern;
//   5. column inverse FFT; 6. row c2r;
//   7. overlap-add inside the tile (src/buffer.cpp:381-560 float order):
//      pixels whose contributing blocks all lie in the tile are rounded,
//      clamped and stored; pixels on a band shared with a neighbouring tile
//      store their partial sums to the seam buffer, finished by k_seam.
#pragma once
#include "nf3d_fft.h"
#include "nf3d_mfma.h"
#include "nf3d_internal.h"

namespace nf3d {

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

// Cover row -> source row of the processed region (interlaced fields are
// stacked, second field bottom-up: src/buffer.cpp:31-63).
__device__ __forceinline__ int cover_to_src_row(const PlaneJob& J, int cy) {
  int ry = reflect101(cy - J.mirh, J.Hc);
  if (J.interlaced) {
    int half = J.Hc >> 1;
    ry = ry < half ? 2 * ry : J.Hc - 1 - 2 * (ry - half);
  }
  return ry;
}

// Output row of a cover interior row (inverse of the above, src/buffer.cpp:99-115).
__device__ __forceinline__ int interior_to_out_row(const PlaneJob& J, int ry) {
  if (!J.interlaced) return ry;
  int half = J.Hc >> 1;
  return ry < half ? 2 * ry : J.Hc - 1 - 2 * (ry - half);
}

// Rounding and clamping of one output pixel (src/buffer.cpp:406: truncation
// toward zero after adding 0.5 and planeBase, then [0, 2^bits-1]; float
// samples clamp to [0,1]).
__device__ __forceinline__ void store_px(const PlaneJob& J, uint8_t* dst, int yo, int x, float v) {
  uint8_t* row = dst + (long long)(yo + J.tc) * J.dst_pitch;
  const int xx = x + J.lc;
  if (J.pix == 4) {
    float o = v < 0.0f ? 0.0f : v;
    o = o > 1.0f ? 1.0f : o;
    reinterpret_cast<float*>(row)[xx] = o;
  } else {
    int q = (int)v;
    q = q < 0 ? 0 : q;
    q = q > J.maxval ? J.maxval : q;
    if (J.pix == 1) row[xx] = (uint8_t)q;
    else reinterpret_cast<uint16_t*>(row)[xx] = (uint16_t)q;
  }
}

// ---------------------------------------------------------------- spectral
// The scalar parameters of the spectral pass, as values (k_run copies them
// out of the launch record once per frame).
struct SpecParams {
  int use_degrid, use_pattern;
  float lowlimit, sharpen, smin_n, smax_n, dehalo, ht2n, sigma2d, kratio2;
  __device__ explicit SpecParams(const PlaneJob& J)
      : use_degrid(J.use_degrid), use_pattern(J.use_pattern), lowlimit(J.lowlimit), sharpen(J.sharpen),
        smin_n(J.smin_n), smax_n(J.smax_n), dehalo(J.dehalo), ht2n(J.ht2n), sigma2d(J.sigma2d), kratio2(J.kratio2) {}
};

// Per-frequency table values of one element, loaded together with the
// spectra so a batch of elements waits on memory once.
struct Tab {
  float2 g;     // gridsample
  float sig;    // pattern (2D or 3D) or the scalar noise power
  float wsh;    // sharpen window
  float wdh;    // dehalo window
};

// `poff`: the block's element offset when the pattern is per block (learned
// PSD), else 0.
template <bool PAT3D>
__device__ __forceinline__ Tab load_tab(const PlaneJob& J, int k, bool sharp, size_t poff) {
  Tab t;
  t.g = J.use_degrid ? J.grid[k] : make_float2(0.0f, 0.0f);
  t.sig = J.use_pattern ? (PAT3D ? J.pattern3d[poff + k] : J.pattern2d[poff + k]) : J.sigma2;
  t.wsh = sharp ? J.wsharpen[k] : 0.0f;
  t.wdh = sharp ? J.wdehalo[k] : 0.0f;
  return t;
}

// Gains of the spectral pass use the hardware reciprocal and square root
// (v_rcp_f32 / v_sqrt_f32, ~1 ulp) and fused multiply-adds: a gain is a
// smooth function of the power, so a few-ulp gain change moves an output
// pixel by ~1e-7 of its value, far below the FP32 transform error the parity
// bar (against the FP64 evaluation of the same formulas) already allows.
__device__ __forceinline__ float fast_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fmaf_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
// |c|^2 + 1e-15 (the reference's psd)
__device__ __forceinline__ float power(float re, float im) { return fmaf_(re, re, fmaf_(im, im, 1.0e-15f)); }

// Factors of src/code_impl/Sharpen.cpp:102-105.
template <class PJ>
__device__ __forceinline__ float sharpen_factor(const PJ& J, float psd, float wsh) {
  float a = J.sharpen * wsh;
  float q = fast_div(psd * J.smax_n, (psd + J.smin_n) * (psd + J.smax_n));
  return fmaf_(a, __builtin_amdgcn_sqrtf(q), 1.0f);
}
template <class PJ>
__device__ __forceinline__ float dehalo_factor(const PJ& J, float psd, float wdh) {
  float s = psd + J.ht2n;
  return fast_div(s, fmaf_(J.dehalo * wdh, psd, s));
}

// Wiener gain of one bin, max((psd - sig) / psd, lowlimit), as
// max(1 - sig / psd, lowlimit) (src/code_impl/Apply.cpp:13-235).
__device__ __forceinline__ float wiener_gain(float re, float im, float sig, float lowlimit) {
  const float f = fmaf_(-sig, __builtin_amdgcn_rcpf(power(re, im)), 1.0f);
  return __builtin_fmaxf(f, lowlimit);  // f is never NaN (psd >= 1e-15)
}

// Apply2D (src/code_impl/Apply.cpp:13-65): one element.
__device__ __forceinline__ float2 apply2d_elem(const PlaneJob& J, float2 x, const Tab& T, float gf) {
  float g0 = 0.0f, g1 = 0.0f;
  if (J.use_degrid) {
    g0 = gf * T.g.x;
    g1 = gf * T.g.y;
  }
  float cr = x.x - g0, ci = x.y - g1;
  float psd = power(cr, ci);
  float f = fmaf_(-T.sig, __builtin_amdgcn_rcpf(psd), 1.0f);
  f = __builtin_fmaxf(f, J.lowlimit);
  if (!J.use_pattern) {
    if (J.dehalo == 0.0f) f *= sharpen_factor(J, psd, T.wsh);
    else if (J.sharpen == 0.0f) f *= dehalo_factor(J, psd, T.wdh);
    else f *= sharpen_factor(J, psd, T.wsh) * dehalo_factor(J, psd, T.wdh);
  }
  return make_float2(fmaf_(cr, f, g0), fmaf_(ci, f, g1));
}

// Sharpen (src/code_impl/Sharpen.cpp:13-133): one element; caller checks
// that sharpen or dehalo is non-zero.
__device__ __forceinline__ float2 sharpen_elem(const PlaneJob& J, float2 x, const Tab& T, float gf) {
  float g0 = 0.0f, g1 = 0.0f;
  if (J.use_degrid) {
    g0 = gf * T.g.x;
    g1 = gf * T.g.y;
  }
  float cr = x.x - g0, ci = x.y - g1;
  float psd = power(cr, ci);
  float f;
  if (J.dehalo == 0.0f) f = sharpen_factor(J, psd, T.wsh);
  else if (J.sharpen == 0.0f) f = dehalo_factor(J, psd, T.wdh);
  else f = sharpen_factor(J, psd, T.wsh) * dehalo_factor(J, psd, T.wdh);
  return make_float2(fmaf_(cr, f, g0), fmaf_(ci, f, g1));
}

// Apply3D2..3D5 (src/code_impl/Apply.cpp:67-235); v[0..BT-1] are the frames
// oldest first.
template <int BT, class PJ>
__device__ __forceinline__ float2 apply3d_elem(const PJ& J, const float2 (&v)[5], const Tab& T, float gf) {
  const float sig = T.sig;
  const float ll = J.lowlimit;
  float g0 = 0.0f, g1 = 0.0f;
  if (J.use_degrid) {
    g0 = gf * T.g.x * (float)BT;
    g1 = gf * T.g.y * (float)BT;
  }
  // the temporal DFT of the reference's formulas, the Wiener gain of each
  // bin, and the inverse at the current frame as one fused sum per part
  float2 r;
  if constexpr (BT == 2) {
    float2 p = v[0], q = v[1];
    float f0r = q.x + p.x - g0, f0i = q.y + p.y - g1;
    float f1r = q.x - p.x, f1i = q.y - p.y;
    const float w0 = wiener_gain(f0r, f0i, sig, ll), w1 = wiener_gain(f1r, f1i, sig, ll);
    r.x = fmaf_(f0r, w0, fmaf_(f1r, w1, g0)) * 0.5f;
    r.y = fmaf_(f0i, w0, fmaf_(f1i, w1, g1)) * 0.5f;
  } else if constexpr (BT == 3) {
    constexpr float s120 = 0.86602540378443864676372317075294f;
    constexpr float third = 1.0f / 3.0f;
    float2 p = v[0], q = v[1], n = v[2];
    float pnr = p.x + n.x, pni = p.y + n.y;
    float fcr = q.x + pnr - g0, fci = q.y + pni - g1;
    float di = s120 * (p.y - n.y);
    float dr = s120 * (n.x - p.x);
    float mr = fmaf_(-0.5f, pnr, q.x), mi = fmaf_(-0.5f, pni, q.y);
    float fpr = mr + di, fnr = mr - di;
    float fpi = mi + dr, fni = mi - dr;
    const float wc = wiener_gain(fcr, fci, sig, ll), wp = wiener_gain(fpr, fpi, sig, ll),
                wn = wiener_gain(fnr, fni, sig, ll);
    r.x = fmaf_(fcr, wc, fmaf_(fpr, wp, fmaf_(fnr, wn, g0))) * third;
    r.y = fmaf_(fci, wc, fmaf_(fpi, wp, fmaf_(fni, wn, g1))) * third;
  } else if constexpr (BT == 4) {
    float2 p2 = v[0], p = v[1], q = v[2], n = v[3];
    float fp2r = (q.x + p2.x) - (p.x + n.x);
    float fp2i = (q.y + p2.y) - (p.y + n.y);
    float fpr = (q.x - p2.x) + (p.y - n.y);
    float fpi = (q.y - p2.y) - (p.x - n.x);
    float fcr = (q.x + p2.x) + (p.x + n.x) - g0;
    float fci = (q.y + p2.y) + (p.y + n.y) - g1;
    float fnr = (q.x - p2.x) - (p.y - n.y);
    float fni = (q.y - p2.y) + (p.x - n.x);
    const float w2 = wiener_gain(fp2r, fp2i, sig, ll), wp = wiener_gain(fpr, fpi, sig, ll),
                wc = wiener_gain(fcr, fci, sig, ll), wn = wiener_gain(fnr, fni, sig, ll);
    r.x = fmaf_(fp2r, w2, fmaf_(fpr, wp, fmaf_(fcr, wc, fmaf_(fnr, wn, g0)))) * 0.25f;
    r.y = fmaf_(fp2i, w2, fmaf_(fpi, wp, fmaf_(fci, wc, fmaf_(fni, wn, g1)))) * 0.25f;
  } else {
    constexpr float s72 = 0.95105651629515357211643933337938f;
    constexpr float c72 = 0.30901699437494742410229341718282f;
    constexpr float s144 = 0.58778525229247312916870595463907f;
    constexpr float c144 = -0.80901699437494742410229341718282f;
    float2 p2 = v[0], p = v[1], q = v[2], n = v[3], n2 = v[4];
    float sum = fmaf_(p2.x + n2.x, c72, fmaf_(p.x + n.x, c144, q.x));
    float dif = fmaf_(n2.y - p2.y, s72, (p.y - n.y) * s144);
    float fp2r = sum + dif, fn2r = sum - dif;
    sum = fmaf_(p2.y + n2.y, c72, fmaf_(p.y + n.y, c144, q.y));
    dif = fmaf_(p2.x - n2.x, s72, (n.x - p.x) * s144);
    float fp2i = sum + dif, fn2i = sum - dif;
    sum = fmaf_(p2.x + n2.x, c144, fmaf_(p.x + n.x, c72, q.x));
    dif = fmaf_(p2.y - n2.y, s144, (p.y - n.y) * s72);
    float fpr = sum + dif, fnr = sum - dif;
    sum = fmaf_(p2.y + n2.y, c144, fmaf_(p.y + n.y, c72, q.y));
    dif = fmaf_(n2.x - p2.x, s144, (n.x - p.x) * s72);
    float fpi = sum + dif, fni = sum - dif;
    float fcr = (p2.x + p.x) + q.x + (n.x + n2.x) - g0;
    float fci = (p2.y + p.y) + q.y + (n.y + n2.y) - g1;
    const float w2 = wiener_gain(fp2r, fp2i, sig, ll), wp = wiener_gain(fpr, fpi, sig, ll),
                wc = wiener_gain(fcr, fci, sig, ll), wn = wiener_gain(fnr, fni, sig, ll),
                wn2 = wiener_gain(fn2r, fn2i, sig, ll);
    r.x = fmaf_(fp2r, w2, fmaf_(fpr, wp, fmaf_(fcr, wc, fmaf_(fnr, wn, fmaf_(fn2r, wn2, g0))))) * 0.2f;
    r.y = fmaf_(fp2i, w2, fmaf_(fpi, wp, fmaf_(fci, wc, fmaf_(fni, wn, fmaf_(fn2i, wn2, g1))))) * 0.2f;
  }
  return r;
}

// Kalman (src/code_impl/Kalman.cpp:27-64): one element, state at offset o.
// `st` holds the state loaded for this element: last, covar, covarProcess.
struct KalState {
  float2 last, covar, cproc;
};

template <class PJ>
__device__ __forceinline__ float2 kalman_elem(const PJ& J, float2 cur, const Tab& T, KalState& st) {
  float sigma = J.use_pattern ? (T.sig < 1e-15f ? 1e-15f : T.sig) : J.sigma2d;
  const float2 prev = st.last;
  const float dr = cur.x - prev.x, di = cur.y - prev.y;
  const float thr = sigma * J.kratio2;
  const bool reset = dr * dr > thr || di * di > thr;
  const float sr = st.covar.x + st.cproc.x, si = st.covar.y + st.cproc.y;
  const float gr = fast_div(sr, sr + sigma), gi = fast_div(si, si + sigma);
  float2 out, cv, cp;
  cp.x = reset ? sigma : gr * gr * sigma;
  cp.y = reset ? sigma : gi * gi * sigma;
  cv.x = reset ? sigma : (1 - gr) * sr;
  cv.y = reset ? sigma : (1 - gi) * si;
  out.x = reset ? cur.x : gr * cur.x + (1 - gr) * prev.x;
  out.y = reset ? cur.y : gi * cur.y + (1 - gi) * prev.y;
  st.last = out;
  st.covar = cv;
  st.cproc = cp;
  return out;
}

// The launch argument read in place from the kernarg segment: a by-value copy
// indexed by the runtime plane would be materialised in scratch memory.
__device__ __forceinline__ const TileLaunch& kernarg_launch(const TileLaunch& arg) {
#if defined(__HIP_DEVICE_COMPILE__)
  (void)arg;
  return *(const TileLaunch*)__builtin_amdgcn_kernarg_segment_ptr();
#else
  return arg;
#endif
}

// ------------------------------------------------------------------ shapes
template <int BW, int BH, int NT>
struct TShape {
  static constexpr int PAIRS = BH / 2;
  static constexpr int COLS = BW / 2 + 1;
  static constexpr int TPB = COLS > PAIRS ? COLS : PAIRS;
  static constexpr int BPW = NT / TPB;
  static constexpr int SR = COLS;                  // spectral row stride (float2)
  static constexpr int LSPEC = BH * SR;            // float2 per block
  static constexpr int SYNW = BW + 1;              // synthesised row stride (float)
  static constexpr int LSYN = BH * SYNW;           // float per block
  static constexpr int SA = BW + 1;                // windowed row-pair stride (float2)
  static constexpr int LA = PAIRS * SA;            // float2 per block (gather stage)
  static constexpr int B_SPEC = BPW * LSPEC * 8;
  static constexpr int B_SYN = BPW * LSYN * 4;
  static constexpr int B_RAW = BPW * LA * 8;
  static constexpr int BYTES = B_SPEC > B_SYN ? (B_SPEC > B_RAW ? B_SPEC : B_RAW) : (B_SYN > B_RAW ? B_SYN : B_RAW);
};

// ------------------------------------------------------------------ phases
// Phase 1: the tile's cover region, windowed into per-block row pairs.
// Lanes own tile columns; each pixel is read once (unconditionally: tail lanes
// re-read a valid column, 16 rows per batch so the batch's loads are in flight
// together) and written, times the analysis window in the reference's product
// order (src/buffer.cpp:137-378), to the one or two blocks holding it:
// (wx*wy)*p on the outer corners of the edge blocks, (wy*p)*wx on the other
// windowed columns of the first/last block rows, (wx*p)*wy elsewhere -- all
// of the form (a*p)*b.
template <int PIX, int BW, int BH, int NT, int SA, int LA>
__device__ __forceinline__ void gather_tile(const PlaneJob& J, float2* stage, const long long* rowoff,
                                            const float* winx, const float* winy, int by, int bx0, int nblk,
                                            int RW, int stepx, int t) {
  using T = typename std::conditional<PIX == 1, uint8_t, typename std::conditional<PIX == 2, uint16_t, float>::type>::type;
  const uint8_t* base = J.src;
  const int ow = J.ow, oh = J.oh;
  for (int c0 = 0; c0 < RW; c0 += NT) {
    const bool cv = c0 + t < RW;
    const int c = cv ? c0 + t : RW - 1;
    const int sxc = reflect101(bx0 * stepx + c - J.mirw, J.Wc) + J.lc;
    // the (up to) two blocks holding tile column c: A = the one starting at or
    // before c with c in its body, B = its left neighbour when c is in B's
    // right band
    int bA = c / stepx, jA;
    if (bA >= nblk) bA = nblk - 1;
    jA = c - bA * stepx;
    const int bB = bA - 1, jB = jA + stepx;
    const bool hasB = bA >= 1 && jB < BW;
    auto xinfo = [&](int b, int j, float& wx, int& side, bool& xedge) {
      side = j < ow ? 1 : (j >= BW - ow ? 2 : 0);
      wx = winx[j];
      const int bx = bx0 + b;
      xedge = (bx == 0 && side == 1) || (bx == J.nox - 1 && side == 2);
    };
    float wxA, wxB = 1.0f;
    int sA, sB = 0;
    bool eA, eB = false;
    xinfo(bA, jA, wxA, sA, eA);
    if (hasB) xinfo(bB, jB, wxB, sB, eB);
    static_for<0, (BH + 15) / 16>([&](auto rc) {
      constexpr int r0 = decltype(rc)::value * 16;
      constexpr int NR = BH - r0 < 16 ? BH - r0 : 16;
      float v[NR];
      static_for<0, NR>([&](auto rr) {
        v[rr] = (float)reinterpret_cast<const T*>(base + rowoff[r0 + decltype(rr)::value])[sxc];
      });
      if (cv) {
        static_for<0, NR>([&](auto rr) {
          constexpr int i = r0 + decltype(rr)::value;
          const float p = v[rr] - J.base;
          const bool ytop = i < oh, ybot = i >= BH - oh;
          const bool yw = ytop || ybot;
          const bool yedge = (by == 0 && ytop) || (by == J.noy - 1 && ybot);
          const float wy = winy[i];
          auto windowed = [&](float wx, int side, bool xedge) {
            float a, b;
            if (side != 0 && yw) {
              a = yedge ? (xedge ? wx * wy : wy) : wx;
              b = yedge ? (xedge ? 1.0f : wx) : wy;
            } else {
              a = side != 0 ? wx : (yw ? wy : 1.0f);
              b = 1.0f;
            }
            return (a * p) * b;
          };
          reinterpret_cast<float*>(&stage[bA * LA + (i >> 1) * SA + jA])[i & 1] = windowed(wxA, sA, eA);
          if (hasB) reinterpret_cast<float*>(&stage[bB * LA + (i >> 1) * SA + jB])[i & 1] = windowed(wxB, sB, eB);
        });
      }
    });
  }
}

// Phase 7: overlap-add of the tile (src/buffer.cpp:381-560 float order).
// A lane owns an output column for all rows; the horizontal combination is
// h = sa*wa + sb*wb with the left block's sample sa and the right block's sb
// (a zero weight drops a side exactly: exclusive columns have wa = 0, wb = 1).
// Row classes are uniform across the workgroup.
template <int PIX, int BW, int BH, int NT, int LSYN, int SYNW>
__device__ __forceinline__ void overlap_tile(const PlaneJob& J, uint8_t* dst, float* seam, const float* syn, int by,
                                             int bx0, int nblk, int RW, int stepx, int stepy, int t) {
  using T = typename std::conditional<PIX == 1, uint8_t, typename std::conditional<PIX == 2, uint16_t, float>::type>::type;
  const float rounder = PIX == 4 ? 0.0f : 0.5f;
  const float fbase = (float)J.obase;
  const int cy0 = by * stepy, cx0 = bx0 * stepx;
  for (int x0 = 0; x0 < RW; x0 += NT) {
    const int xl = x0 + t;
    const int x = cx0 + xl - J.mirw;
    const bool xin = xl < RW && x >= 0 && x < J.Wc;
    // x class: 0 exclusive column, 1 band inside the tile, 2 band shared with
    // the left tile (this tile holds its right block), 3 band shared with the
    // right tile (this tile holds its left block)
    int cls = 0, la = 0, lb = 0;
    float wa = 0.0f, wb = 1.0f;
    if (xl < nblk * stepx) {
      const int bl = xl / stepx, j = xl - bl * stepx;
      lb = bl * LSYN + j;
      la = lb;
      if (j < J.ow) {
        wb = J.wsynxl[j];
        if (bl >= 1) {
          cls = 1;
          la = (bl - 1) * LSYN + j + stepx;
          wa = J.wsynxr[j];
        } else {
          cls = 2;
        }
      }
    } else if (xl < RW) {
      const int jj = xl - nblk * stepx;
      cls = 3;
      la = (nblk - 1) * LSYN + stepx + jj;
      lb = la;
      wa = J.wsynxr[jj];
      wb = 0.0f;
    }
    const size_t seam_slot_x = (size_t)(cls == 3 ? 0 : 1) * J.seam_plane;  // column seams
    T* dcol = reinterpret_cast<T*>(dst + (long long)J.tc * J.dst_pitch) + x + J.lc;
    for (int i = 0; i < BH; i++) {
      const int ry = cy0 + i - J.mirh;
      if (ry < 0 || ry >= J.Hc) continue;  // padding rows (uniform)
      const float h = syn[la + i * SYNW] * wa + syn[lb + i * SYNW] * wb;
      const int yo = interior_to_out_row(J, ry);
      if (!xin) continue;
      const size_t so = (size_t)yo * J.Wc + x;
      if (i >= J.oh && i < stepy) {  // rows of this block row only
        if (cls <= 1) {
          const float v = h * J.norm + rounder + fbase;
          T* d = reinterpret_cast<T*>(reinterpret_cast<uint8_t*>(dcol) + (long long)yo * J.dst_pitch);
          if constexpr (PIX == 4) {
            const float o = v < 0.0f ? 0.0f : v;
            *d = o > 1.0f ? 1.0f : o;
          } else {
            int q = (int)v;
            q = q < 0 ? 0 : q;
            *d = (T)(q > J.maxval ? J.maxval : q);
          }
        } else {
          seam[seam_slot_x + so] = h;  // slot 0 = left block's part, slot 1 = right's
        }
      } else {
        const bool ytop = i < J.oh;
        if (cls <= 1) {
          // row seam: slot 0 = upper block row's part, slot 1 = lower's
          const float w = ytop ? J.wsynyl[i] * J.norm : J.wsynyr[i - stepy] * J.norm;
          seam[(size_t)(ytop ? 1 : 0) * J.seam_plane + so] = h * w;
        } else {
          // corner: A upper-left, B upper-right, C lower-left, D lower-right
          const int slot = ytop ? (cls == 3 ? 2 : 3) : (cls == 3 ? 0 : 1);
          seam[(size_t)slot * J.seam_plane + so] = h;
        }
      }
    }
  }
}

// ------------------------------------------------------------------ kernel
// MF: the 2D transforms as DFT-matrix products on the matrix cores (32x32,
// 48x48 and 64x64 blocks; phases 2-3 and 5-6 become one forward and one inverse product chain
// per block, a wavefront per block), the rest unchanged.
template <int BW, int BH, int NT, int MODE, int BT, int ONCHIP, bool MF = false>
__global__ void __launch_bounds__(NT) k_tile(TileLaunch L_arg) {
  const TileLaunch& L = kernarg_launch(L_arg);
  using S = TShape<BW, BH, NT>;
  constexpr int COLS = S::COLS, PAIRS = S::PAIRS, BPW = S::BPW, SR = S::SR, LSPEC = S::LSPEC;
  constexpr int SYNW = S::SYNW, LSYN = S::LSYN, SA = S::SA, LA = S::LA;
  static_assert(BH % 4 == 0 && BW % 4 == 0, "block sides must be multiples of 4");
  constexpr bool PIXELS = MODE != MODE_W3D || ONCHIP;
  __shared__ __align__(16) unsigned char smem[S::BYTES];
  __shared__ float s_gf[BPW], s_gfs[BPW];
  __shared__ float winx[BW], winy[BH];
  __shared__ long long rowoff[BH];  // byte offset of the source row of each tile row
  float2* spec = reinterpret_cast<float2*>(smem);
  float2* stage = reinterpret_cast<float2*>(smem);
  float* syn = reinterpret_cast<float*>(smem);

  // ---- tile of this workgroup
  int g = blockIdx.x;
  int jb = 0;
  if (L.njobs > 1 && g >= L.job[1].wg_begin) jb = 1;
  if (L.njobs > 2 && g >= L.job[2].wg_begin) jb = 2;
  const PlaneJob& J = L.job[jb];
  g -= J.wg_begin;
  const int by = g / J.ntx;
  const int bx0 = (g - by * J.ntx) * BPW;
  const int nblk = min(BPW, J.nox - bx0);
  const int t = threadIdx.x;
  const int stepx = BW - J.ow, stepy = BH - J.oh;
  const int RW = (nblk - 1) * stepx + BW;
  const size_t tile_spec0 = (size_t)(by * J.nox + bx0) * (BH * COLS);  // first block's spectrum
  const int pix = J.pix;

  const int cb = t / COLS, cc = t - (t / COLS) * COLS;
  const bool col_on = cb < nblk;
  const int rb = t / PAIRS, rp = t - (t / PAIRS) * PAIRS;
  const bool row_on = rb < nblk;

  if constexpr (PIXELS) {
    for (int i = t; i < BW; i += NT) winx[i] = J.wanx[i];
    for (int i = t; i < BH; i += NT) {
      winy[i] = J.wany[i];
      rowoff[i] = (long long)(cover_to_src_row(J, by * stepy + i) + J.tc) * J.src_pitch;
    }
    __syncthreads();
    // ---- 1. windowed pixels of the tile, per block
    if (pix == 1) gather_tile<1, BW, BH, NT, SA, LA>(J, stage, rowoff, winx, winy, by, bx0, nblk, RW, stepx, t);
    else if (pix == 2) gather_tile<2, BW, BH, NT, SA, LA>(J, stage, rowoff, winx, winy, by, bx0, nblk, RW, stepx, t);
    else gather_tile<4, BW, BH, NT, SA, LA>(J, stage, rowoff, winx, winy, by, bx0, nblk, RW, stepx, t);
    __syncthreads();
    if constexpr (MF) {
      // ---- 2-3. 2D r2c on the matrix cores, in rounds of one block per
      // wavefront, highest blocks first: spectral block b overlaps only the
      // gather stage of blocks b and b+1, so a round writes over stages its own
      // round or earlier rounds have read (one barrier per round)
      static_assert(BW == BH && (BW == 32 || BW == 48 || BW == 64), "matrix-core DFT: 32x32, 48x48 or 64x64 blocks");
      static_assert(LSPEC >= LA && 2 * LA - LSPEC >= (BPW - 1) * (LSPEC - LA), "spectra overlap stages b, b+1 only");
      constexpr int NW = NT / 64, MB = (BPW + NW - 1) / NW;
      const int wave = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
      const DftTile* T = static_cast<const DftTile*>(J.dft_tab);
      const float sx = J.dft_sx;
#pragma unroll 1
      for (int q = 0; q < MB; q++) {
        const int b = nblk - 1 - (q * NW + wave);  // uniform in the wavefront
        DftIn<BW> din;
        if (b >= 0)
          din = dft_fwd_load<BW>(
              [&](int r, int c) {
                const float2 v = stage[b * LA + (r >> 1) * SA + c];
                return (r & 1) ? v.y : v.x;
              },
              sx, lane);
        __syncthreads();  // the round's stages read
        if (b >= 0)
          dft_fwd_compute<BW>(din, sx, T, lane, [&](int r, int k, float2 zz) { spec[b * LSPEC + r * SR + k] = zz; });
      }
    } else {
      // ---- 2. row r2c, two real rows per lane
      float2 z[BW];
      if (row_on) {
        static_for<0, BW>([&](auto j) { z[j] = stage[rb * LA + rp * SA + decltype(j)::value]; });
      }
      __syncthreads();  // gather stage dead; spectral layout from here
      if (row_on) {
        fft<BW, -1>(z);
        float2* dst = spec + rb * LSPEC + 2 * rp * SR;
        static_for<0, COLS>([&](auto kk) {
          constexpr int k = decltype(kk)::value;
          constexpr int km = (BW - k) % BW;
          const float2 a = z[k], b = z[km];
          dst[k] = make_float2((a.x + b.x) * 0.5f, (a.y - b.y) * 0.5f);
          dst[SR + k] = make_float2((a.y + b.y) * 0.5f, (b.x - a.x) * 0.5f);
        });
      }
    }
    __syncthreads();
    // ---- 3. column FFT (forward only: store the spectrum and stop)
    if constexpr (MODE == MODE_FWD) {
      if (col_on) {
        float2 col[BH];
        static_for<0, BH>([&](auto i) { col[i] = spec[cb * LSPEC + decltype(i)::value * SR + cc]; });
        if constexpr (!MF) fft<BH, -1>(col);
        float2* out = J.spec_out + tile_spec0 + (size_t)cb * (BH * COLS) + cc;
        static_for<0, BH>([&](auto i) { out[decltype(i)::value * COLS] = col[i]; });
      }
      return;
    }
  }

  if constexpr (MODE != MODE_FWD) {
    const bool do_sharp = J.sharpen != 0.0f || J.dehalo != 0.0f;
    const float2* in0 = J.spec_in[0];
    const float2* in1 = J.spec_in[1];
    const float2* in2 = J.spec_in[2];
    const float2* in3 = J.spec_in[3];
    const float2* in4 = J.spec_in[4];
    auto ring = [&](int q) -> const float2* {
      return q == 0 ? in0 : q == 1 ? in1 : q == 2 ? in2 : q == 3 ? in3 : in4;
    };
    // one lane per spectral column: the on-chip frame's column after the
    // forward column FFT, the other frames' columns from the ring
    const size_t colo = tile_spec0 + (size_t)(col_on ? cb : 0) * (BH * COLS) + cc;  // element (0, cc)
    float2 col[BH];
    if constexpr (PIXELS) {
      if (col_on) {
        static_for<0, BH>([&](auto i) { col[i] = spec[cb * LSPEC + decltype(i)::value * SR + cc]; });
        if constexpr (!MF) fft<BH, -1>(col);
        if (J.spec_out) {
          float2* out = J.spec_out + colo;
          static_for<0, BH>([&](auto i) { out[decltype(i)::value * COLS] = col[i]; });
        }
      }
    }
    // ---- 4a. per-block degrid fractions: from the input DC and from the
    // filtered DC (a following sharpen pass corrects with the filtered
    // buffer's DC, code_impl_C.h:64); lane 0 of each block computes both
    if (col_on && cc == 0) {
      const int b = cb;
      const Tab T0 = load_tab<MODE == MODE_W3D>(J, 0, do_sharp, J.pattern_per_block ? colo - cc : 0);
      float gf = 0.0f, dcf;
      if constexpr (MODE == MODE_W3D) {
        float2 v[5];
#pragma unroll
        for (int q = 0; q < 5; q++) v[q] = make_float2(0.0f, 0.0f);
#pragma unroll
        for (int q = 0; q < BT; q++) v[q] = (ONCHIP && q == BT - 1) ? col[0] : ring(q)[colo];
        const float dcin = v[BT / 2].x;  // current frame: index -from = bt/2
        gf = J.use_degrid ? J.degrid * dcin / J.grid[0].x : 0.0f;
        dcf = apply3d_elem<BT>(J, v, T0, gf).x;
      } else if constexpr (MODE == MODE_W2D) {
        gf = J.use_degrid ? J.degrid * col[0].x / J.grid[0].x : 0.0f;
        dcf = apply2d_elem(J, col[0], T0, gf).x;
      } else if constexpr (MODE == MODE_KALMAN) {
        KalState st{J.kal_last[colo], J.kal_covar[colo], J.kal_cproc[colo]};
        dcf = kalman_elem(J, col[0], T0, st).x;
      } else {
        dcf = col[0].x;
      }
      s_gf[b] = gf;
      s_gfs[b] = J.use_degrid ? J.degrid * dcf / J.grid[0].x : 0.0f;
    }
    __syncthreads();
    // ---- 4b. spectral pass down the column; loads issued 4 rows at a time
    if (col_on) {
      const float gf = s_gf[cb], gfs = s_gfs[cb];
      const size_t poff = J.pattern_per_block ? colo - cc : 0;
      static_for<0, BH / 4>([&](auto ic) {
        constexpr int i0 = decltype(ic)::value * 4;
        float2 v[4][5];
        Tab tb[4];
        KalState ks[4];
        static_for<0, 4>([&](auto uu) {
          constexpr int u = decltype(uu)::value, i = i0 + u;
          const size_t o = colo + (size_t)i * COLS;
#pragma unroll
          for (int q = 0; q < 5; q++) v[u][q] = make_float2(0.0f, 0.0f);
          if constexpr (MODE == MODE_W3D) {
#pragma unroll
            for (int q = 0; q < BT; q++) v[u][q] = (ONCHIP && q == BT - 1) ? col[i] : ring(q)[o];
          } else {
            v[u][0] = col[i];
          }
          if constexpr (MODE == MODE_KALMAN) ks[u] = KalState{J.kal_last[o], J.kal_covar[o], J.kal_cproc[o]};
          tb[u] = load_tab<MODE == MODE_W3D>(J, i * COLS + cc, do_sharp, poff);
        });
        static_for<0, 4>([&](auto uu) {
          constexpr int u = decltype(uu)::value, i = i0 + u;
          float2 r;
          if constexpr (MODE == MODE_W3D) {
            r = apply3d_elem<BT>(J, v[u], tb[u], gf);
            if (do_sharp) r = sharpen_elem(J, r, tb[u], gfs);
          } else if constexpr (MODE == MODE_W2D) {
            r = apply2d_elem(J, v[u][0], tb[u], gf);
            if (J.use_pattern && do_sharp) r = sharpen_elem(J, r, tb[u], gfs);
          } else if constexpr (MODE == MODE_KALMAN) {
            r = kalman_elem(J, v[u][0], tb[u], ks[u]);
            const size_t o = colo + (size_t)i * COLS;
            J.kal_last[o] = ks[u].last;
            J.kal_covar[o] = ks[u].covar;
            J.kal_cproc[o] = ks[u].cproc;
            if (do_sharp) r = sharpen_elem(J, r, tb[u], gfs);
          } else {
            r = do_sharp ? sharpen_elem(J, v[u][0], tb[u], gfs) : v[u][0];
            // pattern display (src/helper.cpp:134-169): every block but the
            // pattern block is zeroed
            if (J.keep_choice && by * J.nox + bx0 + cb != J.keep_choice[1] * J.nox + J.keep_choice[0])
              r = make_float2(0.0f, 0.0f);
          }
          col[i] = r;
        });
      });
      // ---- 5. column inverse FFT, back to LDS for the row c2r
      if constexpr (!MF) fft<BH, +1>(col);
      static_for<0, BH>([&](auto i) { spec[cb * LSPEC + decltype(i)::value * SR + cc] = col[i]; });
    }
    __syncthreads();
    if constexpr (MF) {
      // ---- 5-6. 2D c2r on the matrix cores (DC and Nyquist imaginary parts
      // dropped by the tables), in rounds of one block per wavefront, lowest
      // blocks first: synthesised block b overlaps only the spectra of blocks
      // b-1 and b
      static_assert(LSYN <= 2 * LSPEC && 2 * LSPEC - LSYN <= 2 * LSPEC / (BPW > 1 ? BPW - 1 : 1),
                    "synthesised rows overlap spectra b-1, b only");
      constexpr int NW = NT / 64, MB = (BPW + NW - 1) / NW;
      const int wave = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
      const DftTile* T = static_cast<const DftTile*>(J.dft_tab);
      const float sz = J.dft_sz;
#pragma unroll 1
      for (int q = 0; q < MB; q++) {
        const int b = q * NW + wave;  // uniform in the wavefront
        DftSpec<BW> dsp;
        if (b < nblk) dsp = dft_inv_load<BW>([&](int r, int k) { return spec[b * LSPEC + r * SR + k]; }, sz, lane);
        __syncthreads();  // the round's spectra read
        if (b < nblk)
          dft_inv_compute<BW>(dsp, sz, T, lane, [&](int r, int c0, f4 x) {
            float* d = syn + b * LSYN + r * SYNW + c0;
            d[0] = x[0];
            d[1] = x[1];
            d[2] = x[2];
            d[3] = x[3];
          });
      }
      __syncthreads();
    } else {
    // ---- 6. row c2r (Hermitian rows packed in pairs; DC and Nyquist imag dropped)
    {
      float2 z[BW];
      if (row_on) {
        const float2* src = spec + rb * LSPEC + 2 * rp * SR;
        static_for<0, COLS>([&](auto kk) {
          constexpr int k = decltype(kk)::value;
          float2 a = src[k], b = src[SR + k];
          if constexpr (k == 0 || 2 * k == BW) {
            a.y = 0.0f;
            b.y = 0.0f;
          }
          z[k] = make_float2(a.x - b.y, a.y + b.x);
          if constexpr (k != 0 && 2 * k != BW) z[BW - k] = make_float2(a.x + b.y, b.x - a.y);
        });
      }
      __syncthreads();  // spectral layout dead; synthesised rows next
      if (row_on) {
        fft<BW, +1>(z);
        float* d0 = syn + rb * LSYN + 2 * rp * SYNW;
        static_for<0, BW>([&](auto j) {
          d0[decltype(j)::value] = z[j].x;
          d0[SYNW + decltype(j)::value] = z[j].y;
        });
      }
      __syncthreads();
    }
    }
    // ---- 7. overlap-add inside the tile; seam partials to the seam buffer
    if (pix == 1) overlap_tile<1, BW, BH, NT, LSYN, SYNW>(J, J.dst, J.seam, syn, by, bx0, nblk, RW, stepx, stepy, t);
    else if (pix == 2) overlap_tile<2, BW, BH, NT, LSYN, SYNW>(J, J.dst, J.seam, syn, by, bx0, nblk, RW, stepx, stepy, t);
    else overlap_tile<4, BW, BH, NT, LSYN, SYNW>(J, J.dst, J.seam, syn, by, bx0, nblk, RW, stepx, stepy, t);
  }
}

// One seam pixel: a band shared by two or four tiles, finished from the
// partial sums in the reference's float order (src/buffer.cpp:445-477).
template <int BW, int BH, int NT>
__device__ __forceinline__ void seam_pixel(const PlaneJob& J, const float* sm, uint8_t* dst, int yo, int x) {
  constexpr int BPW = TShape<BW, BH, NT>::BPW;
  const int stepx = BW - J.ow, stepy = BH - J.oh;
  int ry = yo;
  if (J.interlaced) ry = (yo & 1) == 0 ? (yo >> 1) : (J.Hc >> 1) + ((J.Hc - 1 - yo) >> 1);
  const int cy = ry + J.mirh, r = cy / stepy, i = cy - r * stepy;
  const int cx = x + J.mirw, c = cx / stepx, j = cx - c * stepx;
  const bool yband = i < J.oh && r >= 1;
  const bool xseam = j < J.ow && c >= 1 && (c % BPW) == 0;
  if (!yband && !xseam) return;
  const size_t o = (size_t)yo * J.Wc + x;
  const size_t P = J.seam_plane;
  const float rounder = J.pix == 4 ? 0.0f : 0.5f;
  const float fbase = (float)J.obase;
  float v;
  if (yband && xseam) {
    const float wyr = J.wsynyr[i] * J.norm, wyl = J.wsynyl[i] * J.norm;
    const float A = sm[o], B = sm[P + o], C = sm[2 * P + o], D = sm[3 * P + o];
    v = (A + B) * wyr + (C + D) * wyl + rounder + fbase;
  } else if (yband) {
    v = sm[o] + sm[P + o] + rounder + fbase;
  } else {
    v = (sm[o] + sm[P + o]) * J.norm + rounder + fbase;
  }
  store_px(J, dst, yo, x, v);
}

template <int BW, int BH, int NT>
__global__ void __launch_bounds__(256) k_seam(TileLaunch L_arg) {
  const TileLaunch& L = kernarg_launch(L_arg);
  int g = blockIdx.y;
  int jb = 0;
  if (L.njobs > 1 && g >= L.job[1].seam_row_begin) jb = 1;
  if (L.njobs > 2 && g >= L.job[2].seam_row_begin) jb = 2;
  const PlaneJob& J = L.job[jb];
  const int yo = g - J.seam_row_begin;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (yo >= J.Hc || x >= J.Wc) return;
  seam_pixel<BW, BH, NT>(J, J.seam, J.dst, yo, x);
}

template <int BW, int BH, bool MF>
hipError_t launch_tile_modes(int mode, int bt, int onchip, const TileLaunch& L, hipStream_t s) {
  constexpr int NT = 128;
  dim3 grid(L.total_wg), block(NT);
  switch (mode) {
    case MODE_FWD: hipLaunchKernelGGL((k_tile<BW, BH, NT, MODE_FWD, 0, 0, MF>), grid, block, 0, s, L); break;
    case MODE_W2D: hipLaunchKernelGGL((k_tile<BW, BH, NT, MODE_W2D, 0, 0, MF>), grid, block, 0, s, L); break;
    case MODE_KALMAN: hipLaunchKernelGGL((k_tile<BW, BH, NT, MODE_KALMAN, 0, 0, MF>), grid, block, 0, s, L); break;
    case MODE_SHARPEN: hipLaunchKernelGGL((k_tile<BW, BH, NT, MODE_SHARPEN, 0, 0, MF>), grid, block, 0, s, L); break;
    case MODE_W3D:
#define NF3D_W3D(B)                                                                                   \
  case B:                                                                                             \
    if (onchip) hipLaunchKernelGGL((k_tile<BW, BH, NT, MODE_W3D, B, 1, MF>), grid, block, 0, s, L);   \
    else hipLaunchKernelGGL((k_tile<BW, BH, NT, MODE_W3D, B, 0, MF>), grid, block, 0, s, L);          \
    break;
      switch (bt) {
        NF3D_W3D(2)
        NF3D_W3D(3)
        NF3D_W3D(4)
        NF3D_W3D(5)
        default: return hipErrorInvalidValue;
      }
#undef NF3D_W3D
      break;
    default: return hipErrorInvalidValue;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || mode == MODE_FWD) return e;
  dim3 sgrid((L.max_wc + 255) / 256, L.total_seam_rows);
  hipLaunchKernelGGL((k_seam<BW, BH, NT>), sgrid, dim3(256), 0, s, L);
  return hipGetLastError();
}

// mf: the matrix-core DFT variant (32x32, 48x48 and 64x64 blocks)
template <int BW, int BH>
hipError_t launch_tile_shape(int mode, int bt, int onchip, int mf, const TileLaunch& L, hipStream_t s) {
  if (!mf) return launch_tile_modes<BW, BH, false>(mode, bt, onchip, L, s);
  if constexpr (BW == BH && (BW == 32 || BW == 48 || BW == 64)) {
    for (int k = 0; k < L.njobs; k++)
      if (!L.job[k].dft_tab) return hipErrorInvalidValue;
    return launch_tile_modes<BW, BH, true>(mode, bt, onchip, L, s);
  }
  return hipErrorInvalidValue;
}

template <int BW, int BH>
constexpr int tile_bpw() {
  return TShape<BW, BH, 128>::BPW;
}

}  // namespace nf3d
