// Frame-loop kernel of the denoiser: k_run<BW,BH,NT,MODE,BT,SHP,MF>.
//
// A workgroup owns one tile (up to BPW horizontally adjacent blocks of one
// block row, the tile of k_tile) and filters it for a run of consecutive
// frames, so one launch covers many frames and no workgroup waits for the
// others between frames.  Per frame the phases are those of k_tile (gather,
// row r2c, column FFT, spectral pass, column inverse, row c2r, overlap-add);
// what changes is where the data waits are:
//   - the pixels of the next frame are loaded into registers before the
//     current frame's overlap-add, in flight during it and across the barrier;
//   - the spectral pass runs down each column in chunks of rows: the ring
//     spectra / Kalman state of chunk k+DR (DR up to 3) and the table rows of
//     chunk k+1 are loaded while chunk k is computed; the first chunks are
//     loaded before the forward column FFT;
//   - the gains are specialised at compile time on the sharpen / dehalo case
//     (SHP) and degrid is branch-free (a zero fraction when it is off).
// Wiener 3D: the newest frame of each window (frame m + (BT-1)/2) is
// transformed on chip and stored to ring slot (frame % BT); the BT-1 older
// spectra are the ones the same lanes stored in earlier iterations (or the
// launches before), so the loop needs no synchronisation beyond the
// workgroup.  Kalman state likewise stays with its tile.  Seam partials go to
// a per-frame seam buffer that k_seam_run finishes after the launch.  With MF
// the row and column transforms are the matrix-core DFT products of
// nf3d_mfma.h (32x32 blocks).
#pragma once
#include "nf3d_tile.h"

namespace nf3d {

__device__ __forceinline__ const RunBatch& kernarg_batch(const RunBatch& arg) {
#if defined(__HIP_DEVICE_COMPILE__)
  (void)arg;
  return *(const RunBatch*)__builtin_amdgcn_kernarg_segment_ptr();
#else
  return arg;
#endif
}

// The same record through an offset the compiler cannot see through (zero at
// run time): loads of launch fields inside a loop stay in the loop instead of
// being hoisted into registers that live across it.
__device__ __forceinline__ const RunBatch& kernarg_batch_opaque(const RunBatch& arg) {
#if defined(__HIP_DEVICE_COMPILE__)
  (void)arg;
  int zero = 0;
  asm volatile("" : "+s"(zero));
  return *(const RunBatch*)((const char*)__builtin_amdgcn_kernarg_segment_ptr() + zero);
#else
  return arg;
#endif
}

// The job of a workgroup: the last one whose first workgroup it is past.
__device__ __forceinline__ int batch_job(const RunBatch& B, int wg) {
  int jb = 0;
  for (int k = 1; k < B.njobs; k++)
    if (wg >= B.job[k].wg_begin) jb = k;
  return jb;
}

// Wavefront of a lane index, as a scalar value.
__device__ __forceinline__ int wave_of(int t) { return __builtin_amdgcn_readfirstlane(t >> 6); }

template <int I, typename X>
__device__ __forceinline__ X& pick(X& a, X& b) {
  if constexpr (I == 0) return a;
  else return b;
}

// ------------------------------------------------------------------ gather
// Rows handled together by the gather's window read-ahead and the
// overlap-add's LDS reads: a divisor of the rows of a wavefront (8, else 6, 4
// or 2), so no chunk runs past them.
template <int RH>
constexpr int row_chunk() {
  return RH % 8 == 0 ? 8 : RH % 6 == 0 ? 6 : RH % 4 == 0 ? 4 : 2;
}

// Work split of the gather and the overlap-add: each wavefront owns RH = BH /
// (NT / 64) consecutive tile rows, and walks the tile's columns in passes of
// 64 (lane l: column l + 64 q), so no wavefront runs a pass in which most of
// its lanes have no column.
//
// Raw samples of one lane's columns (up to NQ passes) in this wavefront's RH
// rows (rowoff: their source row offsets), kept as bits so no use waits on
// them before the windowing pass.  Every pass is written (zeros past the
// tile's passes) so no old values stay live around the frame loop.
template <int PIX, int RH, int NQ>
__device__ __forceinline__ void gather_issue(const uint8_t* src, const int* rowoff, const int (&sx)[NQ], int npass,
                                             uint32_t (&raw)[NQ][RH]) {
  using T = typename std::conditional<PIX == 1, uint8_t, typename std::conditional<PIX == 2, uint16_t, uint32_t>::type>::type;
  static_for<0, NQ>([&](auto qq) {
    constexpr int q = decltype(qq)::value;
    if (q == 0 || q < npass) {  // uniform
      static_for<0, RH>([&](auto rr) {
        const T* row = reinterpret_cast<const T*>(src + rowoff[decltype(rr)::value]);
        raw[q][rr] = (uint32_t)row[sx[q]];
      });
    } else {
      static_for<0, RH>([&](auto rr) { raw[q][rr] = 0u; });
    }
  });
}

// Windowing of the loaded columns into per-block row pairs: the product
// order of gather_tile ((wx*wy)*p on the outer corners of the edge blocks,
// (wy*p)*wx on the other windowed columns of the first/last block rows,
// (wx*p)*wy elsewhere; src/buffer.cpp:137-378).  A column held by one block
// writes its sample twice to the same place, so no lane branches per row.
template <int PIX, int BW, int BH, int NT, int SA, int LA, int RH, int NQ>
__device__ __forceinline__ void gather_window(const PlaneJob& J, float2* stage, const uint32_t (&raw)[NQ][RH],
                                              const float* winx, const float* winy, int by, int bx0, int nblk,
                                              int RW, int stepx, int t) {
  // launch fields as values: one load each, not one per row
  const int ow = J.ow, oh = J.oh, nox = J.nox, noy = J.noy;
  const float base = J.base;
  // lane in the wavefront; its first row (scalar: the row branches below are
  // uniform per wavefront and compile to scalar branches, not exec masks)
  const int l = t & 63, r0 = wave_of(t) * RH;
  constexpr int CW = row_chunk<RH>();  // window rows read ahead together
  static_for<0, NQ>([&](auto pp) {
    constexpr int p = decltype(pp)::value;
    if (p > 0 && p * 64 >= RW) return;  // uniform
    const int c0 = l + p * 64;
    const bool cv = c0 < RW;
    const int c = cv ? c0 : RW - 1;
    int bA = c / stepx;
    if (bA >= nblk) bA = nblk - 1;
    const int jA = c - bA * stepx;
    const bool hasB = bA >= 1 && jA + stepx < BW;
    const int bB = hasB ? bA - 1 : bA, jB = hasB ? jA + stepx : jA;
    auto xinfo = [&](int b, int j, float& wx, int& side, bool& xedge) {
      side = j < ow ? 1 : (j >= BW - ow ? 2 : 0);
      wx = winx[j];
      const int bx = bx0 + b;
      xedge = (bx == 0 && side == 1) || (bx == nox - 1 && side == 2);
    };
    float wxA, wxB;
    int sA, sB;
    bool eA, eB;
    xinfo(bA, jA, wxA, sA, eA);
    xinfo(bB, jB, wxB, sB, eB);
    float* dA = reinterpret_cast<float*>(&stage[bA * LA + jA]);
    float* dB = reinterpret_cast<float*>(&stage[bB * LA + jB]);
    if (!cv) return;  // tail lanes: nothing to store
    // per-lane factor of the rows outside the vertical bands: wx on a
    // windowed column, else 1 ((1*p)*1 = p and (wx*p)*1 = wx*p exactly)
    const float aA = sA != 0 ? wxA : 1.0f, aB = sB != 0 ? wxB : 1.0f;
    const bool top_edge = by == 0, bot_edge = by == noy - 1;  // uniform
    float wyc[CW];  // window values of the current CW rows, read ahead of their use
    static_for<0, RH>([&](auto rr) {
      constexpr int u = decltype(rr)::value;
      const int i = r0 + u;  // tile row (uniform)
      if constexpr (u % CW == 0) static_for<0, CW>([&](auto uu) { wyc[decltype(uu)::value] = winy[i + decltype(uu)::value]; });
      float v;
      if constexpr (PIX == 4) v = __uint_as_float(raw[p][u]);
      else v = (float)raw[p][u];
      const float px = v - base;
      const bool ytop = i < (BH / 2);                // which band, if any (uniform)
      const bool yw = ytop ? i < oh : i >= BH - oh;  // uniform
      const int off = (i >> 1) * SA * 2 + (u & 1);   // r0 is even
      if (!yw) {
        dA[off] = aA * px;
        dB[off] = aB * px;
      } else {
        const float wy = wyc[u % CW];
        const bool yedge = ytop ? top_edge : bot_edge;  // uniform
        auto windowed = [&](float wx, int side, bool xedge) {
          float a, b;
          if (!yedge) {
            a = side != 0 ? wx : wy;
            b = side != 0 ? wy : 1.0f;
          } else {
            a = side != 0 ? (xedge ? wx * wy : wy) : wy;
            b = side != 0 ? (xedge ? 1.0f : wx) : 1.0f;
          }
          return (a * px) * b;
        };
        dA[off] = windowed(wxA, sA, eA);
        dB[off] = windowed(wxB, sB, eB);
      }
    });
  });
}

// Phase 7 for k_run: overlap-add of the tile in the reference's float order
// (src/buffer.cpp:381-560), as overlap_tile; lanes compute every row and only
// the stores are predicated.  The synthesis windows come from LDS: a global
// load here would wait (vmcnt counts stores too) for every store issued
// before it.
template <int PIX, int BW, int BH, int NT, int LSYN, int SYNW>
__device__ __forceinline__ void overlap_run(const PlaneJob& J, uint8_t* dst, float* seam, const float* syn,
                                            const float* wsyn, int by, int bx0, int nblk, int RW, int stepx,
                                            int stepy, int t) {
  // wsyn: synthesis windows in LDS, [left x (BW) | right x (BW) | top y (BH) |
  // bottom y (BH)]; one base with computed indices (selecting between table
  // pointers would put the pointers in scratch memory)
  constexpr int XL = 0, XR = BW, YL = 2 * BW, YR = 2 * BW + BH;
  using T = typename std::conditional<PIX == 1, uint8_t, typename std::conditional<PIX == 2, uint16_t, float>::type>::type;
  const float rounder = PIX == 4 ? 0.0f : 0.5f;
  // launch fields as values: one load each, not one per row
  const float fbase = (float)J.obase, norm = J.norm;
  const int ow = J.ow, oh = J.oh, mirw = J.mirw, mirh = J.mirh, Wc = J.Wc, Hc = J.Hc, maxval = J.maxval;
  const int interlaced = J.interlaced, half = J.Hc >> 1;
  const long long dpitch = J.dst_pitch;
  const int cy0 = by * stepy, cx0 = bx0 * stepx;
  const size_t P = (size_t)J.seam_plane;
  uint8_t* dbase = dst + (long long)J.tc * dpitch + (long long)J.lc * sizeof(T);
  constexpr int NW = NT / 64, RH = BH / NW, NQ = (TShape<BW, BH, 128>::BPW * BW + 63) / 64;
  constexpr int OC = row_chunk<RH>();  // rows per chunk
  const int l = t & 63, r0 = wave_of(t) * RH;  // lane in the wavefront; its first row (scalar)
  static_for<0, NQ>([&](auto pp) {
    constexpr int p = decltype(pp)::value;
    if (p > 0 && p * 64 >= RW) return;  // uniform
    const int xl = l + p * 64;
    const int x = cx0 + xl - mirw;
    const bool xin = xl < RW && x >= 0 && x < Wc;
    // x class: 0 exclusive column, 1 band inside the tile, 2 band shared with
    // the left tile (this tile holds its right block), 3 band shared with the
    // right tile (this tile holds its left block)
    int cls = 0, la = 0, lb = 0;
    float wa = 0.0f, wb = 1.0f;
    if (xl < nblk * stepx) {
      const int bl = xl / stepx, j = xl - bl * stepx;
      lb = bl * LSYN + j;
      la = lb;
      if (j < ow) {
        wb = wsyn[XL + j];
        if (bl >= 1) {
          cls = 1;
          la = (bl - 1) * LSYN + j + stepx;
          wa = wsyn[XR + j];
        } else {
          cls = 2;
        }
      }
    } else if (xl < RW) {
      const int jj = xl - nblk * stepx;
      cls = 3;
      la = (nblk - 1) * LSYN + stepx + jj;
      lb = la;
      wa = wsyn[XR + jj];
      wb = 0.0f;
    }
    const bool own = xin && cls <= 1;   // pixel finished in this tile (body rows)
    const bool shr = xin && cls >= 2;   // column seam (body rows)
    // 32-bit element offsets from the uniform plane / seam bases (planes and
    // seam buffers are < 2^31 elements: checked by the host)
    const int slot_x = (cls == 3 ? 0 : 1) * (int)P;
    const int xc = xin ? x : 0;
    const int pitch_el = (int)(dpitch / (long long)sizeof(T));
    T* dpl = reinterpret_cast<T*>(dbase);
    // this wavefront's rows in chunks of OC: the chunk's LDS reads are issued
    // together, ahead of the per-row (uniform) branches
#pragma unroll 1
    for (int i0 = r0; i0 < r0 + RH; i0 += OC) {
      float h[OC], w[OC];
      static_for<0, OC>([&](auto uu) {
        constexpr int u = decltype(uu)::value;
        const int i = i0 + u;
        h[u] = syn[la + i * SYNW] * wa + syn[lb + i * SYNW] * wb;
        w[u] = wsyn[i < oh ? YL + i : (i >= stepy ? YR + i - stepy : YL)];
      });
      static_for<0, OC>([&](auto uu) {
        constexpr int u = decltype(uu)::value;
        const int i = i0 + u;
        const int ry = cy0 + i - mirh;
        if (ry >= 0 && ry < Hc) {  // not a padding row (uniform)
          // output row (interlaced fields stacked, second field bottom-up)
          const int yo = !interlaced ? ry : (ry < half ? 2 * ry : Hc - 1 - 2 * (ry - half));
          const int so = yo * Wc + xc;
          if (i >= oh && i < stepy) {  // rows of this block row only (uniform)
            const float v = h[u] * norm + rounder + fbase;
            T q;
            if constexpr (PIX == 4) {
              const float o = v < 0.0f ? 0.0f : v;
              q = o > 1.0f ? 1.0f : o;
            } else {
              int qi = (int)v;
              qi = qi < 0 ? 0 : qi;
              q = (T)(qi > maxval ? maxval : qi);
            }
            if (own) dpl[(unsigned)(yo * pitch_el + xc)] = q;
            if (shr) seam[(unsigned)(slot_x + so)] = h[u];  // slot 0 = left block's part, slot 1 = right's
          } else {
            const bool ytop = i < oh;
            // row seam: slot 0 = upper block row's part, slot 1 = lower's
            // (times its window); corner: A upper-left, B upper-right, C
            // lower-left, D lower-right
            const int slot = cls <= 1 ? (ytop ? 1 : 0) : (ytop ? (cls == 3 ? 2 : 3) : (cls == 3 ? 0 : 1));
            const float val = cls <= 1 ? h[u] * (w[u] * norm) : h[u];
            if (xin) seam[(unsigned)(slot * (int)P + so)] = val;
          }
        }
      });
    }
  });
}

// ------------------------------------------- gather / overlap-add by column
// 256-lane tiles: lane (block b, column j) reads block b's column j of every
// row straight from the plane (reflect-101 addressing) and writes it,
// windowed, into the block's row pairs -- no per-pass column mapping, one
// store per sample.  The window product is (p * wx) * wy for every sample:
// the same value as the region-by-region order of gather_window up to one
// rounding (the parity bar is against the FP64 product).
template <int PIX, int BW, int BH, int SA, int LA>
__device__ __forceinline__ void gather_cols(const PlaneJob& J, float2* stage, const uint8_t* src, const int* rowoff,
                                            const float* winx, const float* winy, int bx0, int nblk, int stepx,
                                            int t) {
  using T = typename std::conditional<PIX == 1, uint8_t, typename std::conditional<PIX == 2, uint16_t, float>::type>::type;
  const int b = t / BW, j = t - b * BW;
  if (b >= nblk) return;  // lanes past the tile's blocks
  const int sx = reflect101((bx0 + b) * stepx + j - J.mirw, J.Wc) + J.lc;
  const float wx = winx[j], base = J.base;
  // 32-bit byte offsets from the uniform plane base: one add per sample
  // (rows are < 2^31 bytes from the base: checked by the host)
  const int sxb = sx * (int)sizeof(T);
  T v[BH];
  static_for<0, BH>([&](auto ii) {
    constexpr int i = decltype(ii)::value;
    v[i] = *reinterpret_cast<const T*>(src + (unsigned)(rowoff[i] + sxb));
  });
  float* d = reinterpret_cast<float*>(&stage[b * LA + j]);
  static_for<0, BH>([&](auto ii) {
    constexpr int i = decltype(ii)::value;
    const float px = (float)v[i] - base;
    d[(i >> 1) * SA * 2 + (i & 1)] = (px * wx) * winy[i];
  });
}

// Overlap-add with one lane per tile column, down all BH rows of the tile's
// block row: the column's block(s) and synthesis weights are worked out once
// per frame, then each row is one or two LDS reads and its store -- the
// reference's float order of overlap_run (src/buffer.cpp:381-560).
template <int PIX, int BW, int BH, int LSYN, int SYNW>
__device__ __forceinline__ void overlap_cols(const PlaneJob& J, uint8_t* dst, float* seam, const float* syn,
                                             const float* wsyn, int by, int bx0, int nblk, int RW, int stepx,
                                             int stepy, int t) {
  constexpr int XL = 0, XR = BW, YL = 2 * BW, YR = 2 * BW + BH;
  using T = typename std::conditional<PIX == 1, uint8_t, typename std::conditional<PIX == 2, uint16_t, float>::type>::type;
  const int xl = t;
  if (xl >= RW) return;  // lanes past the tile's columns
  const float rounder = PIX == 4 ? 0.0f : 0.5f;
  const float fbase = (float)J.obase, norm = J.norm;
  const int ow = J.ow, oh = J.oh, mirw = J.mirw, mirh = J.mirh, Wc = J.Wc, Hc = J.Hc, maxval = J.maxval;
  const int interlaced = J.interlaced, half = J.Hc >> 1;
  const long long dpitch = J.dst_pitch;
  const int cy0 = by * stepy, cx0 = bx0 * stepx;
  const int P = (int)J.seam_plane;
  T* dpl = reinterpret_cast<T*>(dst + (long long)J.tc * dpitch + (long long)J.lc * sizeof(T));
  const int pitch_el = (int)(dpitch / (long long)sizeof(T));
  const int x = cx0 + xl - mirw;
  const bool xin = x >= 0 && x < Wc;
  // x class: 0 exclusive column, 1 band inside the tile, 2 band shared with
  // the left tile (this tile holds its right block), 3 band shared with the
  // right tile (this tile holds its left block)
  int cls = 0, la = 0, lb = 0;
  float wa = 0.0f, wb = 1.0f;
  if (xl < nblk * stepx) {
    const int bl = xl / stepx, j = xl - bl * stepx;
    lb = bl * LSYN + j;
    la = lb;
    if (j < ow) {
      wb = wsyn[XL + j];
      if (bl >= 1) {
        cls = 1;
        la = (bl - 1) * LSYN + j + stepx;
        wa = wsyn[XR + j];
      } else {
        cls = 2;
      }
    }
  } else {
    const int jj = xl - nblk * stepx;
    cls = 3;
    la = (nblk - 1) * LSYN + stepx + jj;
    lb = la;
    wa = wsyn[XR + jj];
    wb = 0.0f;
  }
  const bool own = xin && cls <= 1, shr = xin && cls >= 2;
  const int slot_x = (cls == 3 ? 0 : 1) * P;
  const int xc = xin ? x : 0;
#pragma unroll 8
  for (int i = 0; i < BH; i++) {
    const int ry = cy0 + i - mirh;
    if (ry < 0 || ry >= Hc) continue;  // padding row (uniform)
    const float h = syn[la + i * SYNW] * wa + syn[lb + i * SYNW] * wb;
    const int yo = !interlaced ? ry : (ry < half ? 2 * ry : Hc - 1 - 2 * (ry - half));
    const int so = yo * Wc + xc;
    if (i >= oh && i < stepy) {  // rows of this block row only (uniform)
      const float v = h * norm + rounder + fbase;
      T q;
      if constexpr (PIX == 4) {
        const float o = v < 0.0f ? 0.0f : v;
        q = o > 1.0f ? 1.0f : o;
      } else {
        int qi = (int)v;
        qi = qi < 0 ? 0 : qi;
        q = (T)(qi > maxval ? maxval : qi);
      }
      if (own) dpl[(unsigned)(yo * pitch_el + xc)] = q;
      if (shr) seam[(unsigned)(slot_x + so)] = h;  // slot 0 = left block's part, slot 1 = right's
    } else {
      const bool ytop = i < oh;
      const float w = wsyn[ytop ? YL + i : YR + i - stepy];
      const int slot = cls <= 1 ? (ytop ? 1 : 0) : (ytop ? (cls == 3 ? 2 : 3) : (cls == 3 ? 0 : 1));
      const float val = cls <= 1 ? h * (w * norm) : h;
      if (xin) seam[(unsigned)(slot * P + so)] = val;
    }
  }
}

// ------------------------------------------------------------ spectral pass
// Sharpen / dehalo factor of one element (src/code_impl/Sharpen.cpp:102-105);
// for SHP_GENERAL the sharpen factor is exactly 1 when sharpen == 0, so the
// product equals the reference's dehalo-only factor.
template <int SHP>
__device__ __forceinline__ float shp_factor(const SpecParams& J, float psd, float wsh, float wdh) {
  if constexpr (SHP == SHP_SHARPEN) return sharpen_factor(J, psd, wsh);
  else return sharpen_factor(J, psd, wsh) * dehalo_factor(J, psd, wdh);
}

// Sharpen pass of one element (src/code_impl/Sharpen.cpp:13-133); degrid
// fraction zero when degrid is off.
template <int SHP>
__device__ __forceinline__ float2 sharpen_s(const SpecParams& J, float2 x, const Tab& T, float gf) {
  const float g0 = gf * T.g.x, g1 = gf * T.g.y;
  const float cr = x.x - g0, ci = x.y - g1;
  const float psd = power(cr, ci);
  const float f = shp_factor<SHP>(J, psd, T.wsh, T.wdh);
  return make_float2(fmaf_(cr, f, g0), fmaf_(ci, f, g1));
}

// Apply2D (src/code_impl/Apply.cpp:13-65) with the sharpen factor folded in
// when no pattern is used.
template <int SHP>
__device__ __forceinline__ float2 apply2d_s(const SpecParams& J, float2 x, const Tab& T, float gf) {
  const float g0 = gf * T.g.x, g1 = gf * T.g.y;
  const float cr = x.x - g0, ci = x.y - g1;
  const float psd = power(cr, ci);
  float f = fmaf_(-T.sig, __builtin_amdgcn_rcpf(psd), 1.0f);
  f = __builtin_fmaxf(f, J.lowlimit);  // f is never NaN (psd >= 1e-15)
  if constexpr (SHP != SHP_NONE) {
    if (!J.use_pattern) f *= shp_factor<SHP>(J, psd, T.wsh, T.wdh);
  }
  return make_float2(fmaf_(cr, f, g0), fmaf_(ci, f, g1));
}

// One chunk of CH rows of a lane's column: the ring spectra / Kalman state
// (HBM, loaded several chunks ahead) and the table values (cache-resident,
// one chunk ahead).
template <int MODE, int NQ, int CH>
struct RingBuf {
  float2 v[NQ > 0 ? NQ : 1][CH];
  KalState ks[MODE == MODE_KALMAN ? CH : 1];
};
template <int SHP, int CH>
struct TabBuf {
  float2 g[CH];
  float sig[CH];
  float wsh[SHP != SHP_NONE ? CH : 1];
  float wdh[SHP == SHP_GENERAL ? CH : 1];
};

// ------------------------------------------------------- split column FFT
// The value of the other lane of a lane pair (lanes 2c, 2c+1): a DPP
// quad permutation [1,0,3,2], one VALU instruction per 32-bit value.
__device__ __forceinline__ float pair_swap(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float2 pair_swap(float2 v) { return make_float2(pair_swap(v.x), pair_swap(v.y)); }

// Forward BH-point column DFT split over a lane pair: lane h holds the
// BH/2-point DFT of the column's rows h, h+2, h+4, ... (E for h = 0, O for
// h = 1).  X[k] = E[k] + w^k O[k] and X[k + BH/2] = E[k] - w^k O[k]
// (w = exp(-2 pi i / BH)): lane 1 twiddles, the pair exchanges, lane 0 keeps
// X[k] and lane 1 X[k + BH/2], k < BH/2.  `sg` is +1 on lane 0, -1 on lane 1
// (fma by +-1 is an exact add / subtract).
template <int BH>
__device__ __forceinline__ void split_combine_fwd(float2 (&c)[BH / 2], int h, float sg) {
  static_for<0, BH / 2>([&](auto kk) {
    constexpr int k = decltype(kk)::value;
    float2 v = c[k];
    if constexpr (k != 0) {
      constexpr float wr = (float)ct::cos2pi(k, BH), wi = (float)(-ct::sin2pi(k, BH));
      const float2 tw = cmul(v, wr, wi);
      v = h ? tw : v;
    }
    const float2 p = pair_swap(v);
    c[k] = make_float2(__builtin_fmaf(sg, v.x, p.x), __builtin_fmaf(sg, v.y, p.y));
  });
}

// Inverse of the above (unnormalised, FFTW c2r convention): lane 0 holds
// X[k], lane 1 X[k + BH/2]; the even output rows are the BH/2-point inverse
// DFT of X[k] + X[k + BH/2], the odd rows that of w^-k (X[k] - X[k + BH/2]).
// Lane h ends with the inputs of its BH/2-point inverse transform.
template <int BH>
__device__ __forceinline__ void split_combine_inv(float2 (&c)[BH / 2], int h, float sg) {
  static_for<0, BH / 2>([&](auto kk) {
    constexpr int k = decltype(kk)::value;
    const float2 v = c[k];
    const float2 p = pair_swap(v);
    float2 d = make_float2(__builtin_fmaf(sg, v.x, p.x), __builtin_fmaf(sg, v.y, p.y));
    if constexpr (k != 0) {
      constexpr float wr = (float)ct::cos2pi(k, BH), wi = (float)ct::sin2pi(k, BH);
      const float2 tw = cmul(d, wr, wi);
      d = h ? tw : d;
    }
    c[k] = d;
  });
}

// ------------------------------------------------------------------ kernel
// Register arrays (row and column transforms, chunk buffers) are computed by
// every lane, idle lanes on a duplicate of a valid column (indices clamped to
// the tile), and only stores are predicated: a register written under a lane
// mask keeps its old value in the other lanes, which would hold it live around
// the frame loop.
//
// MF: the 2D transforms as DFT-matrix products on the matrix cores (32x32
// and 48x48 blocks, nf3d_mfma.h), a wavefront per block in turn, in place of the row
// and column FFTs; the spectral pass is unchanged.
//
// NT = 256 (FFT form): the tile of the 128-lane geometry (BPW blocks) with
// twice the lanes.  Gather and overlap-add split the tile's rows over four
// wavefronts; the row transforms use the first BPW * BH/2 lanes (the other
// wavefronts skip them); each spectral column has a lane pair, which splits
// its transform in two half-length transforms (split_combine_*) and runs the
// spectral pass over half of the column's rows each.
//
// Waves per SIMD the register budget is cut for (256 lanes: LDS allows three
// workgroups per CU at 48x48): three, or two for the variants whose spectral
// pass holds more per element (four or five frames, the dehalo factor, the
// Kalman state with a sharpen pass, 64-row columns), which would otherwise
// spill.  128 lanes (the matrix-core form): two at 32x32, one at 48x48 (256
// VGPRs and the AGPRs, no spills).
template <int BH, int NT, int MODE, int BT, int SHP, bool MF>
constexpr int run_waves_per_simd() {
  if constexpr (MF && BH >= 48) return 1;
  else if constexpr (NT != 256) return 2;
  else if constexpr (BH >= 64 || (MODE == MODE_W3D && (BT >= 4 || SHP == SHP_GENERAL)) ||
                     (MODE == MODE_KALMAN && SHP != SHP_NONE))
    return 2;
  else return 3;
}

template <int BW, int BH, int NT, int MODE, int BT, int SHP, bool MF = false>
__global__ void __launch_bounds__(NT, (run_waves_per_simd<BH, NT, MODE, BT, SHP, MF>())) k_run(RunBatch B_arg) {
  const RunBatch& B0 = kernarg_batch(B_arg);
  const int jb = batch_job(B0, blockIdx.x);
  const RunLaunch& L0 = B0.job[jb];
  using S = TShape<BW, BH, 128>;
  constexpr bool SPLIT = NT == 256;
  static_assert(NT == 128 || (NT == 256 && !MF), "256 lanes: the FFT form");
  constexpr int CR = SPLIT ? BH / 2 : BH;  // spectral rows per column lane
  constexpr int COLS = S::COLS, PAIRS = S::PAIRS, BPW = S::BPW, SR = S::SR, LSPEC = S::LSPEC;
  constexpr int SYNW = S::SYNW, LSYN = S::LSYN, SA = S::SA, LA = S::LA;
  constexpr int FROM = MODE == MODE_W3D ? -(BT / 2) : 0;
  constexpr int TO = MODE == MODE_W3D ? (BT - 1) / 2 : 0;
  constexpr int NQ = MODE == MODE_W3D ? BT - 1 : 0;  // ring frames per element
  constexpr int CH = (MODE == MODE_W3D) ? 4 : 2;  // rows per chunk (two chunks of loads in flight)
  constexpr int NCH = CR / CH;
  static_assert(CR % CH == 0 && BH % 8 == 0 && BW % 4 == 0, "block shape");
  // MODE_FWD: the forward transforms only, frame m0 + i stored to ring slot
  // (m0 + i) % BT -- the Wiener-3D window frames a range starts with, so
  // every spectrum in the ring comes from this kernel's arithmetic
  // chunks of ring rows in flight ahead of the one being computed
  // (256 lanes: half the rows per lane and a register budget of 168 for
  // three wavefronts per SIMD)
  constexpr int DR = MODE == MODE_W3D ? (SPLIT ? (NQ <= 2 ? 2 : 1) : (NQ <= 2 ? 3 : NQ == 3 ? 2 : 1)) : 1;
  using RBuf = RingBuf<MODE, NQ, CH>;
  using TBuf = TabBuf<SHP, CH>;
  __shared__ __align__(16) unsigned char smem[S::BYTES];
  __shared__ float s_gf[BPW], s_gfs[BPW];
  __shared__ __align__(16) float winx[BW];
  __shared__ __align__(16) float winy[BH];
  __shared__ __align__(16) float wsyn[2 * BW + 2 * BH];  // synthesis windows: left, right, top, bottom bands
  // byte offset of the source row of each tile row (< 2^31: checked by the host)
  __shared__ __align__(16) int rowoff[BH];
  float2* spec = reinterpret_cast<float2*>(smem);
  float2* stage = reinterpret_cast<float2*>(smem);
  float* syn = reinterpret_cast<float*>(smem);

  // ---- tile and frames of this workgroup
  const int g = blockIdx.x - L0.wg_begin;
  const int chunk = g / L0.ntiles;
  const int tile = g - chunk * L0.ntiles;
  const int i0 = chunk * L0.fpw;
  const int i1 = min(L0.nframes, i0 + L0.fpw);
  const int by = tile / L0.J.ntx;
  const int bx0 = (tile - by * L0.J.ntx) * BPW;
  const int nblk = min(BPW, L0.J.nox - bx0);
  const int stepx = BW - L0.J.ow, stepy = BH - L0.J.oh;
  const int RW = (nblk - 1) * stepx + BW;
  // gather / overlap-add split: RH rows per wavefront, up to GQ column passes
  constexpr int RH = BH / (NT / 64), GQ = (BPW * BW + 63) / 64;
  static_assert(NT % 64 == 0 && BH % (NT / 64) == 0 && RH % 2 == 0, "row split");
  const int npass = (RW + 63) >> 6;

  {
    const PlaneJob& J = L0.J;
    const int t = threadIdx.x;
    for (int i = t; i < BW; i += NT) winx[i] = J.wanx[i];
    for (int i = t; i < BH; i += NT) {
      winy[i] = J.wany[i];
      rowoff[i] = (int)((long long)(cover_to_src_row(J, by * stepy + i) + J.tc) * J.src_pitch);
    }
    for (int i = t; i < J.ow; i += NT) {
      wsyn[i] = J.wsynxl[i];
      wsyn[BW + i] = J.wsynxr[i];
    }
    for (int i = t; i < J.oh; i += NT) {
      wsyn[2 * BW + i] = J.wsynyl[i];
      wsyn[2 * BW + BH + i] = J.wsynyr[i];
    }
  }
  __syncthreads();

  // pixels of a frame's on-chip transform (this wavefront's rows), loaded
  // one frame ahead
  uint32_t raw[GQ][RH];
  auto issue = [&](const PlaneJob& J, const uint8_t* src, const int* ro, int t) {
    const int l = t & 63;
    int sx[GQ];
    static_for<0, GQ>([&](auto qq) {
      constexpr int q = decltype(qq)::value;
      sx[q] = reflect101(bx0 * stepx + min(l + 64 * q, RW - 1) - J.mirw, J.Wc) + J.lc;
    });
    const int* rw = ro + wave_of(t) * RH;
    if (J.pix == 1) gather_issue<1, RH, GQ>(src, rw, sx, npass, raw);
    else if (J.pix == 2) gather_issue<2, RH, GQ>(src, rw, sx, npass, raw);
    else gather_issue<4, RH, GQ>(src, rw, sx, npass, raw);
  };
  if constexpr (!SPLIT) issue(L0.J, L0.src[i0], rowoff, (int)threadIdx.x);

#pragma unroll 1
  for (int fi = i0; fi < i1; fi++) {
    // The lane index, the launch record and the LDS tables are reached
    // through values the compiler cannot see through, so lane addresses and
    // launch fields are recomputed inside the loop instead of hundreds of
    // loop-invariant values being kept live across it.
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    int zo = 0;
    asm volatile("" : "+s"(zo));
    const RunLaunch& L = kernarg_batch_opaque(B_arg).job[jb];
    const PlaneJob& J = L.J;
    const float* wx = winx + zo;
    const float* wy = winy + zo;
    const int* ro = rowoff + zo;
    const int m = L.m0 + fi;
    const int pix = J.pix;
    // column lanes: one per column, or a pair (SPLIT: hh = half, rows
    // hh * CR .. hh * CR + CR - 1 of the spectrum after the forward transform)
    const int cl = SPLIT ? (t >> 1) : t;
    const int hh = SPLIT ? (t & 1) : 0;
    const float sg = hh ? -1.0f : 1.0f;
    const int cbr = cl / COLS, cc = cl - cbr * COLS;
    const bool col_on = cbr < nblk;
    const int cb = col_on ? cbr : nblk - 1;
    // row lanes: one per row pair, or a pair of lanes (SPLIT, half hh)
    const int rl = SPLIT ? (t >> 1) : t;
    const int rbr = rl / PAIRS, rp = rl - rbr * PAIRS;
    const bool row_on = rbr < nblk;
    const int rb = row_on ? rbr : nblk - 1;
    // wavefronts with no row pair of the tile skip the row transforms
    const bool row_wave = wave_of(t) * 64 < BPW * PAIRS * (SPLIT ? 2 : 1);
    const size_t tile_spec0 = (size_t)(by * J.nox + bx0) * (BH * COLS);
    // element (hh * CR, cc) of this lane's column: its first spectral row
    const size_t colo = tile_spec0 + (size_t)cb * (BH * COLS) + (size_t)(hh * CR) * COLS + cc;
    // the pattern: one for the plane, or this block's (learned PSD)
    const float* pat = J.use_pattern ? (MODE == MODE_W3D ? J.pattern3d : J.pattern2d) +
                                           (J.pattern_per_block ? colo - (size_t)(hh * CR) * COLS - cc : 0) +
                                           (size_t)(hh * CR) * COLS
                                     : nullptr;
    const float grid0 = J.grid[0].x;
    const SpecParams SP(J);

    // ---- 1. window the loaded pixels into per-block row pairs
    if constexpr (SPLIT) {
      const uint8_t* fsrc = L.src[fi];
      if (pix == 1) gather_cols<1, BW, BH, SA, LA>(J, stage, fsrc, ro, wx, wy, bx0, nblk, stepx, t);
      else if (pix == 2) gather_cols<2, BW, BH, SA, LA>(J, stage, fsrc, ro, wx, wy, bx0, nblk, stepx, t);
      else gather_cols<4, BW, BH, SA, LA>(J, stage, fsrc, ro, wx, wy, bx0, nblk, stepx, t);
    } else {
      if (pix == 1) gather_window<1, BW, BH, NT, SA, LA, RH, GQ>(J, stage, raw, wx, wy, by, bx0, nblk, RW, stepx, t);
      else if (pix == 2) gather_window<2, BW, BH, NT, SA, LA, RH, GQ>(J, stage, raw, wx, wy, by, bx0, nblk, RW, stepx, t);
      else gather_window<4, BW, BH, NT, SA, LA, RH, GQ>(J, stage, raw, wx, wy, by, bx0, nblk, RW, stepx, t);
    }
    __syncthreads();
    if constexpr (MF) {
      // ---- 2-3. 2D r2c on the matrix cores, in rounds of one block per
      // wavefront, highest blocks first.  Spectral block b overlaps only the
      // gather stage of blocks b and b+1 (static_assert below), so a round
      // writes over stages its own round or earlier rounds have read, and one
      // barrier per round (operands loaded by both wavefronts) is enough.
      static_assert(BW == BH && (BW == 32 || BW == 48 || BW == 64), "matrix-core DFT: 32x32, 48x48 or 64x64 blocks");
      static_assert(LSPEC >= LA && 2 * LA - LSPEC >= (BPW - 1) * (LSPEC - LA), "spectra overlap stages b, b+1 only");
      constexpr int NW = NT / 64, MB = (BPW + NW - 1) / NW;
      const int wave = wave_of(t), lane = t & 63;
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
    } else if (!row_wave) {
      __syncthreads();  // gather stage dead; spectral layout from here
    } else if constexpr (SPLIT) {
      // ---- 2. row r2c, a lane pair per two real rows packed as one complex
      // row Z: the BW-point transform as two BW/2-point transforms of its even
      // and odd columns combined across the pair (lane 0: Z[k], lane 1:
      // Z[k + BW/2]); after an exchange each lane finishes half of the output
      // columns k (lane 0: 0..BW/4, lane 1: BW/4+1..BW/2) from Z[k] and
      // Z[BW - k].
      constexpr int HW = BW / 2, K0 = HW / 2 + 1;
      float2 z[HW];
      static_for<0, HW>([&](auto j) { z[j] = stage[rb * LA + rp * SA + 2 * decltype(j)::value + hh]; });
      __syncthreads();  // gather stage dead; spectral layout from here
      fft<HW, -1>(z);
      split_combine_fwd<BW>(z, hh, sg);
      float2 o[HW];
      static_for<0, HW>([&](auto j) { o[j] = pair_swap(z[j]); });
      // lane 0: low half Z[0..HW) = z, high half = o; lane 1 the other way round
      float2* dst = spec + rb * LSPEC + 2 * rp * SR + hh * K0;
      static_for<0, K0>([&](auto qq) {
        constexpr int q = decltype(qq)::value;
        constexpr int k1 = K0 + q;  // lane 1's column
        const float2 a0 = z[q], b0 = q == 0 ? z[0] : o[(HW - q) % HW];
        const float2 a1 = k1 < HW ? o[k1 < HW ? k1 : 0] : z[0];
        const float2 b1 = z[k1 <= HW ? HW - k1 : 0];
        const float2 a = hh ? a1 : a0, b = hh ? b1 : b0;
        if (row_on && (hh == 0 || k1 <= HW)) {
          dst[q] = make_float2((a.x + b.x) * 0.5f, (a.y - b.y) * 0.5f);
          dst[SR + q] = make_float2((a.y + b.y) * 0.5f, (b.x - a.x) * 0.5f);
        }
      });
    } else {
      // ---- 2. row r2c, two real rows per lane
      float2 z[BW];
      static_for<0, BW>([&](auto j) { z[j] = stage[rb * LA + rp * SA + decltype(j)::value]; });
      __syncthreads();  // gather stage dead; spectral layout from here
      fft<BW, -1>(z);
      if (row_on) {
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

    if constexpr (MODE == MODE_FWD) {
      // ---- 3. forward column transforms to the ring, then the next frame
      float2 fc[CR];
      if constexpr (SPLIT) {
        static_for<0, CR>([&](auto j) { fc[j] = spec[cb * LSPEC + (2 * decltype(j)::value + hh) * SR + cc]; });
      } else {
        static_for<0, CR>([&](auto i) { fc[i] = spec[cb * LSPEC + decltype(i)::value * SR + cc]; });
      }
      __syncthreads();  // spectra read: the next frame's gather may write the stage
      if constexpr (!SPLIT) issue(J, L.src[min(fi + 1, i1 - 1)], ro, t);
      if constexpr (SPLIT) {
        fft<CR, -1>(fc);
        split_combine_fwd<BH>(fc, hh, sg);
      } else if constexpr (!MF) {
        fft<BH, -1>(fc);
      }
      if (col_on) {
        float2* out = L.ring + (size_t)(m % BT) * L.spec_elems + colo;
        static_for<0, CR>([&](auto i) { out[decltype(i)::value * COLS] = fc[i]; });
      }
      continue;
    }
    // ---- 3. column lanes: chunk 0 in flight during the forward column FFT
    const float2* rq[NQ > 0 ? NQ : 1];
    static_for<0, NQ>([&](auto qq) {
      constexpr int q = decltype(qq)::value;
      rq[q] = L.ring + (size_t)((m + FROM + q) % BT) * L.spec_elems;
    });
    auto load_ring = [&](auto kc, RBuf& b) {
      constexpr int k = decltype(kc)::value;
      static_for<0, CH>([&](auto uu) {
        constexpr int u = decltype(uu)::value, i = k * CH + u;
        const size_t o = colo + (size_t)i * COLS;
        static_for<0, NQ>([&](auto qq) { b.v[decltype(qq)::value][u] = rq[decltype(qq)::value][o]; });
        if constexpr (MODE == MODE_KALMAN) b.ks[u] = KalState{J.kal_last[o], J.kal_covar[o], J.kal_cproc[o]};
      });
    };
    const int tab0 = hh * CR * COLS + cc;  // table element of the lane's first row
    auto load_tab = [&](auto kc, TBuf& b) {
      constexpr int k = decltype(kc)::value;
      static_for<0, CH>([&](auto uu) {
        constexpr int u = decltype(uu)::value, i = k * CH + u;
        const int kk = i * COLS + tab0;
        b.g[u] = J.grid[kk];
        if constexpr (SHP != SHP_NONE) b.wsh[u] = J.wsharpen[kk];
        if constexpr (SHP == SHP_GENERAL) b.wdh[u] = J.wdehalo[kk];
      });
      if (pat) {
        static_for<0, CH>([&](auto uu) { b.sig[decltype(uu)::value] = pat[(k * CH + decltype(uu)::value) * COLS + cc]; });
      } else {
        static_for<0, CH>([&](auto uu) { b.sig[decltype(uu)::value] = J.sigma2; });
      }
    };
    auto tab = [&](const TBuf& b, auto uu) {
      constexpr int u = decltype(uu)::value;
      Tab T;
      T.g = b.g[u];
      T.sig = b.sig[u];
      T.wsh = 0.0f;
      T.wdh = 0.0f;
      if constexpr (SHP != SHP_NONE) T.wsh = b.wsh[u];
      if constexpr (SHP == SHP_GENERAL) T.wdh = b.wdh[u];
      return T;
    };
    RBuf rbuf[DR + 1];
    TBuf tbuf[2];
    float2 col[CR];
    static_for<0, (DR < NCH ? DR : NCH)>([&](auto kc) { load_ring(kc, rbuf[decltype(kc)::value]); });
    load_tab(std::integral_constant<int, 0>{}, tbuf[0]);
    if constexpr (SPLIT) {
      // even rows on lane 0, odd rows on lane 1
      static_for<0, CR>([&](auto j) { col[j] = spec[cb * LSPEC + (2 * decltype(j)::value + hh) * SR + cc]; });
      fft<CR, -1>(col);
      split_combine_fwd<BH>(col, hh, sg);
    } else {
      static_for<0, BH>([&](auto i) { col[i] = spec[cb * LSPEC + decltype(i)::value * SR + cc]; });
      if constexpr (!MF) fft<BH, -1>(col);
    }
    if constexpr (MODE == MODE_W3D) {
      if (col_on) {
        float2* out = L.ring + (size_t)((m + TO) % BT) * L.spec_elems + colo;
        static_for<0, CR>([&](auto i) { out[decltype(i)::value * COLS] = col[i]; });
      }
    }
    // ---- 4a. per-block degrid fractions from the input DC and from the
    // filtered DC (code_impl_C.h:64); lane 0 of each block computes both
    if (col_on && cc == 0 && hh == 0) {
      const Tab T0 = tab(tbuf[0], std::integral_constant<int, 0>{});
      float gf = 0.0f, dcf;
      if constexpr (MODE == MODE_W3D) {
        float2 w[5];
#pragma unroll
        for (int q = 0; q < 5; q++) w[q] = make_float2(0.0f, 0.0f);
        static_for<0, NQ>([&](auto qq) { w[decltype(qq)::value] = rbuf[0].v[decltype(qq)::value][0]; });
        w[NQ] = col[0];
        gf = J.use_degrid ? J.degrid * w[BT / 2].x / grid0 : 0.0f;
        dcf = apply3d_elem<BT>(SP, w, T0, gf).x;
      } else if constexpr (MODE == MODE_W2D) {
        gf = J.use_degrid ? J.degrid * col[0].x / grid0 : 0.0f;
        dcf = apply2d_s<SHP>(SP, col[0], T0, gf).x;
      } else if constexpr (MODE == MODE_KALMAN) {
        KalState st = rbuf[0].ks[0];
        dcf = kalman_elem(SP, col[0], T0, st).x;
      } else {
        dcf = col[0].x;
      }
      s_gf[cb] = gf;
      s_gfs[cb] = J.use_degrid ? J.degrid * dcf / grid0 : 0.0f;
    }
    __syncthreads();
    // ---- 4b. spectral pass down the column, chunk k+1 loading under chunk k
    {
      const float gf = s_gf[cb], gfs = s_gfs[cb];
      static_for<0, NCH>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if constexpr (k + DR < NCH) load_ring(std::integral_constant<int, k + DR>{}, rbuf[(k + DR) % (DR + 1)]);
        if constexpr (k + 1 < NCH) load_tab(std::integral_constant<int, k + 1>{}, tbuf[(k + 1) & 1]);
        const RBuf& b = rbuf[k % (DR + 1)];
        const TBuf& tb = tbuf[k & 1];
        static_for<0, CH>([&](auto uu) {
          constexpr int u = decltype(uu)::value, i = k * CH + u;
          const Tab T = tab(tb, uu);
          float2 r;
          if constexpr (MODE == MODE_W3D) {
            float2 w[5];
#pragma unroll
            for (int q = 0; q < 5; q++) w[q] = make_float2(0.0f, 0.0f);
            static_for<0, NQ>([&](auto qq) { w[decltype(qq)::value] = b.v[decltype(qq)::value][u]; });
            w[NQ] = col[i];
            r = apply3d_elem<BT>(SP, w, T, gf);
            if constexpr (SHP != SHP_NONE) r = sharpen_s<SHP>(SP, r, T, gfs);
          } else if constexpr (MODE == MODE_W2D) {
            r = apply2d_s<SHP>(SP, col[i], T, gf);
            if constexpr (SHP != SHP_NONE) {
              const float2 rs = sharpen_s<SHP>(SP, r, T, gfs);
              r = SP.use_pattern ? rs : r;
            }
          } else if constexpr (MODE == MODE_KALMAN) {
            KalState st = b.ks[u];
            r = kalman_elem(SP, col[i], T, st);
            if (col_on) {
              const size_t o = colo + (size_t)i * COLS;
              J.kal_last[o] = st.last;
              J.kal_covar[o] = st.covar;
              J.kal_cproc[o] = st.cproc;
            }
            if constexpr (SHP != SHP_NONE) r = sharpen_s<SHP>(SP, r, T, gfs);
          } else {
            r = sharpen_s<SHP>(SP, col[i], T, gfs);
          }
          col[i] = r;
        });
        // chunk k's results are pinned here (an empty asm that clobbers
        // memory): the compiler neither sinks the gains below the later
        // chunks' loads nor hoists those loads above them, so at most two
        // chunks of loads are in registers at once
        static_for<0, CH>([&](auto uu) {
          float2& cv = col[k * CH + decltype(uu)::value];
          float a = cv.x, b = cv.y;
          asm volatile("" : "+v"(a), "+v"(b) : : "memory");
          cv.x = a;
          cv.y = b;
        });
      });
    }
    // ---- 5. column inverse FFT, back to LDS for the row c2r
    if constexpr (SPLIT) {
      split_combine_inv<BH>(col, hh, sg);
      fft<CR, +1>(col);
      if (col_on) {
        static_for<0, CR>([&](auto j) { spec[cb * LSPEC + (2 * decltype(j)::value + hh) * SR + cc] = col[j]; });
      }
    } else {
      if constexpr (!MF) fft<BH, +1>(col);
      if (col_on) {
        static_for<0, BH>([&](auto i) { spec[cb * LSPEC + decltype(i)::value * SR + cc] = col[i]; });
      }
    }
    __syncthreads();
    if constexpr (MF) {
      // ---- 5-6. 2D c2r on the matrix cores (DC and Nyquist imaginary parts
      // dropped by the tables), in rounds of one block per wavefront, lowest
      // blocks first: synthesised block b overlaps only the spectra of blocks
      // b-1 and b (static_assert below)
      static_assert(LSYN <= 2 * LSPEC && 2 * LSPEC - LSYN <= 2 * LSPEC / (BPW > 1 ? BPW - 1 : 1),
                    "synthesised rows overlap spectra b-1, b only");
      constexpr int NW = NT / 64, MB = (BPW + NW - 1) / NW;
      const int wave = wave_of(t), lane = t & 63;
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
    } else if (!row_wave) {
      __syncthreads();  // spectral layout dead; synthesised rows next
    } else if constexpr (SPLIT) {
      // ---- 6. row c2r, a lane pair per two rows: each lane builds the packed
      // spectrum z (Hermitian rows packed, DC and Nyquist imaginary parts
      // dropped) and reduces it to its half-length inverse transform's input
      // (split_combine_inv: lane 0 z[k] + z[k + BW/2], lane 1
      // w^-k (z[k] - z[k + BW/2])); lane h ends with output columns 2j + h.
      constexpr int HW = BW / 2;
      const float2* src = spec + rb * LSPEC + 2 * rp * SR;
      auto zcol = [&](auto kk) {  // z[k] (k <= HW) from column k
        constexpr int k = decltype(kk)::value;
        float2 a = src[k], b = src[SR + k];
        if constexpr (k == 0 || k == HW) {
          a.y = 0.0f;
          b.y = 0.0f;
        }
        return make_float2(a.x - b.y, a.y + b.x);
      };
      auto zmirror = [&](auto kk) {  // z[BW - k] (0 < k < HW) from column k
        constexpr int k = decltype(kk)::value;
        const float2 a = src[k], b = src[SR + k];
        return make_float2(a.x + b.y, b.x - a.y);
      };
      float2 z[HW];
      static_for<0, HW>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        const float2 zl = zcol(kk);
        float2 zu;
        if constexpr (k == 0) zu = zcol(std::integral_constant<int, HW>{});
        else zu = zmirror(std::integral_constant<int, HW - k>{});
        float2 d = make_float2(__builtin_fmaf(sg, zu.x, zl.x), __builtin_fmaf(sg, zu.y, zl.y));
        if constexpr (k != 0) {
          constexpr float wr = (float)ct::cos2pi(k, BW), wi = (float)ct::sin2pi(k, BW);
          const float2 tw = cmul(d, wr, wi);
          d = hh ? tw : d;
        }
        z[k] = d;
      });
      __syncthreads();  // spectral layout dead; synthesised rows next
      fft<HW, +1>(z);
      if (row_on) {
        float* d0 = syn + rb * LSYN + 2 * rp * SYNW + hh;
        static_for<0, HW>([&](auto j) {
          d0[2 * decltype(j)::value] = z[j].x;
          d0[SYNW + 2 * decltype(j)::value] = z[j].y;
        });
      }
    } else {
      // ---- 6. row c2r (Hermitian rows packed in pairs; DC and Nyquist imag dropped)
      float2 z[BW];
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
      __syncthreads();  // spectral layout dead; synthesised rows next
      fft<BW, +1>(z);
      if (row_on) {
        float* d0 = syn + rb * LSYN + 2 * rp * SYNW;
        static_for<0, BW>([&](auto j) {
          d0[decltype(j)::value] = z[j].x;
          d0[SYNW + decltype(j)::value] = z[j].y;
        });
      }
    }
    __syncthreads();
    // ---- the next frame's pixels, in flight during the overlap-add and
    // across the barrier (the registers of the row transforms are dead here;
    // the last iteration reloads its own frame: an unconditional load keeps
    // the old values from living through the iteration)
    if constexpr (!SPLIT) issue(J, L.src[min(fi + 1, i1 - 1)], ro, t);
    // ---- 7. overlap-add inside the tile; seam partials to this frame's seam buffer
    uint8_t* dst = L.dst[fi];
    float* seam = J.seam + (size_t)fi * 4 * (size_t)J.seam_plane;
    const float* ws = wsyn + zo;
    if constexpr (SPLIT) {
      if (pix == 1) overlap_cols<1, BW, BH, LSYN, SYNW>(J, dst, seam, syn, ws, by, bx0, nblk, RW, stepx, stepy, t);
      else if (pix == 2) overlap_cols<2, BW, BH, LSYN, SYNW>(J, dst, seam, syn, ws, by, bx0, nblk, RW, stepx, stepy, t);
      else overlap_cols<4, BW, BH, LSYN, SYNW>(J, dst, seam, syn, ws, by, bx0, nblk, RW, stepx, stepy, t);
    } else {
      if (pix == 1) overlap_run<1, BW, BH, NT, LSYN, SYNW>(J, dst, seam, syn, ws, by, bx0, nblk, RW, stepx, stepy, t);
      else if (pix == 2) overlap_run<2, BW, BH, NT, LSYN, SYNW>(J, dst, seam, syn, ws, by, bx0, nblk, RW, stepx, stepy, t);
      else overlap_run<4, BW, BH, NT, LSYN, SYNW>(J, dst, seam, syn, ws, by, bx0, nblk, RW, stepx, stepy, t);
    }
    __syncthreads();  // syn read; the next frame's gather writes the stage
  }
}

// n / d for 0 <= n < 2^22 and 1 <= d <= 64 (inv = 1/d rounded): the float
// quotient is within one of the exact one, and one correction each way
// makes it exact.
__device__ __forceinline__ int div_small(int n, int d, float inv) {
  int q = (int)((float)n * inv);
  const int r = n - q * d;
  q += r >= d ? 1 : 0;
  q -= r < 0 ? 1 : 0;
  return q;
}

// Seam pixels of every frame of a k_run launch, finished from the partial
// sums in the reference's float order (src/buffer.cpp:445-477); one workgroup
// per output row and frame.  Rows inside a horizontal band finish the whole
// row: both band partials are loaded unconditionally (8 pixels per lane in
// flight), the corner partials where a vertical seam crosses.  Other rows
// finish only the columns of the vertical seams between tiles.
template <int BW, int BH, int NT>
__global__ void __launch_bounds__(256) k_seam_run(RunBatch B_arg) {
  const RunLaunch& L = kernarg_batch(B_arg).job[blockIdx.z];
  const PlaneJob& J = L.J;
  constexpr int BPW = TShape<BW, BH, NT>::BPW;
  const int yo = blockIdx.x;
  const int fi = blockIdx.y;
  if (yo >= J.Hc || fi >= L.nframes) return;
  const float* sm = J.seam + (size_t)fi * 4 * (size_t)J.seam_plane;
  uint8_t* dst = L.dst[fi];
  const int stepx = BW - J.ow, stepy = BH - J.oh;
  int ry = yo;
  if (J.interlaced) ry = (yo & 1) == 0 ? (yo >> 1) : (J.Hc >> 1) + ((J.Hc - 1 - yo) >> 1);
  const int cy = ry + J.mirh, r = cy / stepy, i = cy - r * stepy;
  const float rounder = J.pix == 4 ? 0.0f : 0.5f;
  const float fbase = (float)J.obase;
  const size_t P = (size_t)J.seam_plane;
  const float* row = sm + (size_t)yo * J.Wc;
  if (i < J.oh && r >= 1) {
    const float wyr = J.wsynyr[i] * J.norm, wyl = J.wsynyl[i] * J.norm;
    const int Wc = J.Wc, mirw = J.mirw, ow = J.ow;
    const float inv = 1.0f / (float)stepx;
    constexpr int U = 8;
    for (int x0 = 0; x0 < Wc; x0 += 256 * U) {
      // every partial of the lane's pixels in flight together: the band
      // partials always, the corner partials where a vertical seam crosses
      float up[U], lo[U], cA[U], cB[U];
      bool cor[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int x = min(x0 + u * 256 + (int)threadIdx.x, Wc - 1);
        const int cx = x + mirw, c = div_small(cx, stepx, inv), j = cx - c * stepx;
        cor[u] = j < ow && c >= 1 && (c % BPW) == 0;
        up[u] = row[x];
        lo[u] = row[P + x];
        cA[u] = 0.0f;
        cB[u] = 0.0f;
        if (cor[u]) {
          cA[u] = row[2 * P + x];
          cB[u] = row[3 * P + x];
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int x = x0 + u * 256 + (int)threadIdx.x;
        if (x >= Wc) break;
        const float A = up[u], B = lo[u], C = cA[u], D = cB[u];
        const float v = cor[u] ? (A + B) * wyr + (C + D) * wyl + rounder + fbase : A + B + rounder + fbase;
        store_px(J, dst, yo, x, v);
      }
    }
  } else {
    const int nseam = (J.nox - 1) / BPW;  // tile boundaries at block columns BPW, 2*BPW, ...
    for (int k = threadIdx.x; k < nseam * J.ow; k += 256) {
      const int c = (k / J.ow + 1) * BPW, j = k - (k / J.ow) * J.ow;
      const int x = c * stepx + j - J.mirw;
      if (x >= 0 && x < J.Wc) {
        const float v = (row[x] + row[P + x]) * J.norm + rounder + fbase;
        store_px(J, dst, yo, x, v);
      }
    }
  }
}

template <int BW, int BH, bool MF>
hipError_t launch_run_modes(int mode, int bt, int shp, const RunBatch& L, hipStream_t s) {
  constexpr int NT = MF ? 128 : 256;
  dim3 grid(L.total_wg), block(NT);
#define NF3D_RUN(M, B, H) hipLaunchKernelGGL((k_run<BW, BH, NT, M, B, H, MF>), grid, block, 0, s, L)
#define NF3D_RUN_SHP(M, B)                                   \
  switch (shp) {                                             \
    case SHP_NONE: NF3D_RUN(M, B, SHP_NONE); break;          \
    case SHP_SHARPEN: NF3D_RUN(M, B, SHP_SHARPEN); break;    \
    case SHP_GENERAL: NF3D_RUN(M, B, SHP_GENERAL); break;    \
    default: return hipErrorInvalidValue;                    \
  }
  switch (mode) {
    case MODE_FWD:
      switch (bt) {
        case 2: NF3D_RUN(MODE_FWD, 2, SHP_NONE); break;
        case 3: NF3D_RUN(MODE_FWD, 3, SHP_NONE); break;
        case 4: NF3D_RUN(MODE_FWD, 4, SHP_NONE); break;
        case 5: NF3D_RUN(MODE_FWD, 5, SHP_NONE); break;
        default: return hipErrorInvalidValue;
      }
      break;
    case MODE_W2D: NF3D_RUN_SHP(MODE_W2D, 0); break;
    case MODE_KALMAN: NF3D_RUN_SHP(MODE_KALMAN, 0); break;
    case MODE_SHARPEN:
      if (shp == SHP_SHARPEN) NF3D_RUN(MODE_SHARPEN, 0, SHP_SHARPEN);
      else if (shp == SHP_GENERAL) NF3D_RUN(MODE_SHARPEN, 0, SHP_GENERAL);
      else return hipErrorInvalidValue;
      break;
    case MODE_W3D:
      switch (bt) {
        case 2: NF3D_RUN_SHP(MODE_W3D, 2); break;
        case 3: NF3D_RUN_SHP(MODE_W3D, 3); break;
        case 4: NF3D_RUN_SHP(MODE_W3D, 4); break;
        case 5: NF3D_RUN_SHP(MODE_W3D, 5); break;
        default: return hipErrorInvalidValue;
      }
      break;
    default: return hipErrorInvalidValue;
  }
#undef NF3D_RUN_SHP
#undef NF3D_RUN
  return hipGetLastError();
}

// mf: the matrix-core DFT form (32x32 and 48x48 blocks).
template <int BW, int BH>
hipError_t launch_run_shape(int mode, int bt, int shp, int mf, const RunBatch& L, hipStream_t s) {
  if (L.njobs < 1 || L.njobs > RUN_MAX_JOBS) return hipErrorInvalidValue;
  int wg = 0;
  for (int k = 0; k < L.njobs; k++) {
    const RunLaunch& R = L.job[k];
    if (R.nframes < 1 || R.nframes > RUN_MAX_FRAMES || R.fpw < 1 || R.wg_begin != wg) return hipErrorInvalidValue;
    if (mf && !R.J.dft_tab) return hipErrorInvalidValue;
    wg += R.ntiles * ((R.nframes + R.fpw - 1) / R.fpw);
  }
  if (wg != L.total_wg) return hipErrorInvalidValue;
  if (mf) {
    // 64x64 blocks in the matrix-core form go through the per-frame kernel
    // (in the frame loop their operands would exceed the register file)
    if constexpr (BW == BH && (BW == 32 || BW == 48)) return launch_run_modes<BW, BH, true>(mode, bt, shp, L, s);
    else return hipErrorInvalidValue;
  }
  return launch_run_modes<BW, BH, false>(mode, bt, shp, L, s);
}

template <int BW, int BH>
hipError_t launch_seam_run_shape(const RunBatch& L, hipStream_t s) {
  if (L.njobs < 1 || L.njobs > RUN_MAX_JOBS || L.max_frames < 1 || L.max_frames > RUN_MAX_FRAMES) return hipErrorInvalidValue;
  dim3 sgrid(L.max_hc, L.max_frames, L.njobs);
  hipLaunchKernelGGL((k_seam_run<BW, BH, 128>), sgrid, dim3(256), 0, s, L);
  return hipGetLastError();
}

}  // namespace nf3d
