/*
 * Test infrastructure only: the oracle.  Used by tests/, by smoke() and by
 * bench.py's CPU-baseline leg as the checker -- never shipped, never the
 * thing measured.
 *
 * An FP64 host restatement of neo_fft3d's per-plane overlapped-block 3D
 * frequency-domain denoiser, from the reference's own sources:
 *   - front end: defaults, planes, crop and frame window  src/fft3d.hpp:74-264
 *   - engine ctor: sigma scaling, validation, geometry, windows,
 *     sharpen/dehalo/pattern tables, Kalman init, gridsample
 *                                                   src/fft3d_engine.h:87-407
 *   - engine GetFrame: btcur edge rule, 2D/3D/Kalman/sharpen-only modes
 *                                                   src/fft3d_engine.h:452-734
 *   - cover padding (reflect-101), interlaced fields src/buffer.cpp:5-116
 *   - windowed block gather                         src/buffer.cpp:118-379
 *   - synthesis overlap-add, rounding, clamping     src/buffer.cpp:381-560
 *   - Wiener 2D / 3D2..3D5 with degrid + sharpen/dehalo
 *                                                   src/code_impl/Apply.cpp:13-235
 *   - Sharpen / dehalo                              src/code_impl/Sharpen.cpp:13-133
 *   - Kalman                                        src/code_impl/Kalman.cpp:13-67
 *   - per-block degrid fraction, lowlimit           src/code_impl/code_impl_C.h:44-92
 *   - pattern helpers                               src/helper.cpp:14-179
 *   - FFT contract (FFTW unnormalised r2c/c2r)      src/fft3d_engine.h:161-188
 *
 * Tables are built exactly as the reference builds them (float arithmetic,
 * glibc cosf and double exp/sqrt promotion) because they are part of the
 * algorithm's definition; every per-pixel and per-frequency computation runs
 * in double.  The temporal transform of Apply3D2..3D5 is written as a general
 * length-bt DFT along t (the reference's hand-expanded formulas are that DFT;
 * checked term by term against src/code_impl/Apply.cpp:67-235).
 *
 * Pinned by tests/test_oracle.py against: the compiled reference (golden
 * vectors in tests/golden/), the closed forms (FFT vs direct DFT, sigma=0
 * identity, constant frames, window partition of unity, numpy.fft along t).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cplx;

/* ------------------------------------------------------------------ FFT -- */

/* exp(-2 pi i k / n), k < n, cached per length (computed in double by libm). */
static const cplx* twiddles(int n) {
  static cplx* cache[1025];
  if (n > 1024) return NULL;
  if (!cache[n]) {
    cplx* t = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    for (int k = 0; k < n; k++) {
      double a = -2.0 * M_PI * (double)k / n;
      t[k].re = cos(a);
      t[k].im = sin(a);
    }
    cache[n] = t;
  }
  return cache[n];
}

static cplx tw(int e, int n, int sign) {
  const cplx* t = twiddles(n);
  cplx w;
  e %= n;
  if (t) { w = t[e]; }
  else { double a = -2.0 * M_PI * (double)e / n; w.re = cos(a); w.im = sin(a); }
  if (sign > 0) w.im = -w.im;
  return w;
}

/* Mixed-radix decimation-in-time complex DFT of length n (any n), sign -1
 * forward, +1 inverse, unnormalised: out[k] = sum_t in[t*stride] w^(sign*t*k). */
static void dft_rec(const cplx* in, int stride, cplx* out, int n, int sign) {
  if (n == 1) { out[0] = in[0]; return; }
  int p = 2;
  while (p * p <= n && n % p) p++;
  if (n % p) p = n; /* prime length: direct */
  int m = n / p;
  if (m == 1) {
    for (int k = 0; k < n; k++) {
      double sr = 0, si = 0;
      for (int t = 0; t < n; t++) {
        cplx w = tw((int)((long)t * k % n), n, sign);
        sr += in[t * stride].re * w.re - in[t * stride].im * w.im;
        si += in[t * stride].re * w.im + in[t * stride].im * w.re;
      }
      out[k].re = sr; out[k].im = si;
    }
    return;
  }
  cplx sub_stack[256];
  cplx* sub = n <= 256 ? sub_stack : (cplx*)malloc(sizeof(cplx) * n);
  for (int r = 0; r < p; r++) dft_rec(in + r * stride, stride * p, sub + r * m, m, sign);
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) {
      int kk = k + q * m;
      double sr = 0, si = 0;
      for (int r = 0; r < p; r++) {
        cplx w = tw((int)((long)r * kk % n), n, sign);
        const cplx v = sub[r * m + k];
        sr += v.re * w.re - v.im * w.im;
        si += v.re * w.im + v.im * w.re;
      }
      out[kk].re = sr; out[kk].im = si;
    }
  }
  if (sub != sub_stack) free(sub);
}

void nf3d_oracle_dft(const double* in, double* out, int n, int sign) {
  dft_rec((const cplx*)in, 1, (cplx*)out, n, sign);
}

/* 2D r2c of one bh x bw real block -> bh x (bw/2+1) complex (FFTW layout). */
static void r2c2d(const double* x, cplx* X, int bw, int bh, cplx* tmp) {
  int ow = bw / 2 + 1;
  cplx* row = tmp;           /* bw */
  cplx* rowo = tmp + bw;     /* bw */
  cplx* col = tmp + 2 * bw;  /* bh */
  cplx* colo = col + bh;     /* bh */
  cplx* mid = colo + bh;     /* bh * ow */
  for (int i = 0; i < bh; i++) {
    for (int j = 0; j < bw; j++) { row[j].re = x[i * bw + j]; row[j].im = 0; }
    dft_rec(row, 1, rowo, bw, -1);
    for (int k = 0; k < ow; k++) mid[i * ow + k] = rowo[k];
  }
  for (int k = 0; k < ow; k++) {
    for (int i = 0; i < bh; i++) col[i] = mid[i * ow + k];
    dft_rec(col, 1, colo, bh, -1);
    for (int i = 0; i < bh; i++) X[i * ow + k] = colo[i];
  }
}

/* 2D c2r (FFTW semantics: column inverse DFTs, then a Hermitian row c2r that
 * uses only the real parts of the DC and Nyquist terms), unnormalised. */
static void c2r2d(const cplx* X, double* x, int bw, int bh, cplx* tmp) {
  int ow = bw / 2 + 1;
  cplx* col = tmp;
  cplx* colo = tmp + bh;
  cplx* mid = colo + bh;          /* bh * ow */
  cplx* row = mid + bh * ow;      /* bw */
  cplx* rowo = row + bw;          /* bw */
  for (int k = 0; k < ow; k++) {
    for (int i = 0; i < bh; i++) col[i] = X[i * ow + k];
    dft_rec(col, 1, colo, bh, +1);
    for (int i = 0; i < bh; i++) mid[i * ow + k] = colo[i];
  }
  for (int i = 0; i < bh; i++) {
    for (int k = 0; k < ow; k++) row[k] = mid[i * ow + k];
    row[0].im = 0;
    if (bw % 2 == 0) row[bw / 2].im = 0;
    for (int k = ow; k < bw; k++) { row[k].re = row[bw - k].re; row[k].im = -row[bw - k].im; }
    dft_rec(row, 1, rowo, bw, +1);
    for (int j = 0; j < bw; j++) x[i * bw + j] = rowo[j].re;
  }
}

/* ------------------------------------------------------------- params -- */

typedef struct {
  float sigma, beta;
  int bw, bh, bt, ow, oh;
  float kratio, sharpen, scutoff, svr, smin, smax;
  int measure, interlaced, wintype, pframe, px, py, pshow;
  float pcutoff, pfactor, sigma2, sigma3, sigma4, degrid, dehalo, hr, ht;
  int l, t, r, b, opt, ncpu, mt;
  int planes[4], nplanes_list; /* VS "planes" */
  int preg_x, preg_y;           /* "pregion=rx[,ry]": per-region patterns (no reference counterpart) */
} oparams;

typedef struct {
  int family;      /* 0 gray, 1 yuv, 2 rgb */
  int is_float;
  int bits, bytes;
  int ssw, ssh;
  int nplanes;
} oformat;

/* ------------------------------------------------------------- engine -- */

typedef struct {
  oparams ep;    /* scaled copy, as FFT3DEngine keeps it */
  oformat fmt;
  int chroma;
  int frames;
  int framewidth, frameheight;   /* plane size */
  int lc, tc, rc, bc;            /* crop in plane units */
  int Wc, Hc;                    /* copied (processed) region */
  int nox, noy, nb, outwidth, mirw, mirh, coverwidth, coverheight;
  float norm;
  float *wanxl, *wanxr, *wanyl, *wanyr, *wsynxl, *wsynxr, *wsynyl, *wsynyr;
  float *wsharpen, *wdehalo, *pwin;   /* [bh*outwidth] */
  double *pattern2d, *pattern3d;      /* [bh*outwidth] */
  int pattern_set, pattern3d_init;
  int pshow_flat;                      /* pshow has set the windows to 1 */
  double* bpsd;                        /* [nb][bh*outwidth] per-block noise PSD of the next frame, or NULL */
  int preg_x, preg_y;                  /* per-region pattern estimation: region size in blocks (0: off) */
  cplx* grid;                          /* [bh*outwidth] */
  float sigmaSquaredNoiseNormed2D, sigmaSquaredSharpenMinNormed, sigmaSquaredSharpenMaxNormed, ht2n;
  /* Kalman state */
  cplx *outLast, *covar, *covarProcess;
  /* spectra cache (results-neutral), one per frame */
  cplx** spec;
  /* scratch */
  double* blk;
  cplx* tmp;
} oengine;

static void* xcalloc(size_t n, size_t s) { void* p = calloc(n ? n : 1, s); if (!p) abort(); return p; }

/* Reference windows, src/fft3d_engine.h:204-283 (float arithmetic). */
static void make_windows(oengine* e) {
  int ow = e->ep.ow, oh = e->ep.oh;
  float pi = 3.1415926535897932384626433832795f;
  e->wanxl = xcalloc(ow, 4); e->wanxr = xcalloc(ow, 4);
  e->wanyl = xcalloc(oh, 4); e->wanyr = xcalloc(oh, 4);
  e->wsynxl = xcalloc(ow, 4); e->wsynxr = xcalloc(ow, 4);
  e->wsynyl = xcalloc(oh, 4); e->wsynyr = xcalloc(oh, 4);
  if (e->ep.wintype == 0) {
    for (int i = 0; i < ow; i++) {
      e->wanxl[i] = cosf(pi * ((float)(i - ow) + 0.5f) / (float)(ow * 2));
      e->wanxr[i] = cosf(pi * ((float)i + 0.5f) / (float)(ow * 2));
      e->wsynxl[i] = e->wanxl[i]; e->wsynxr[i] = e->wanxr[i];
    }
    for (int i = 0; i < oh; i++) {
      e->wanyl[i] = cosf(pi * ((float)(i - oh) + 0.5f) / (float)(oh * 2));
      e->wanyr[i] = cosf(pi * ((float)i + 0.5f) / (float)(oh * 2));
      e->wsynyl[i] = e->wanyl[i]; e->wsynyr[i] = e->wanyr[i];
    }
  } else if (e->ep.wintype == 1) {
    for (int i = 0; i < ow; i++) {
      e->wanxl[i] = (float)sqrt((double)cosf(pi * ((float)(i - ow) + 0.5f) / (float)(ow * 2)));
      e->wanxr[i] = (float)sqrt((double)cosf(pi * ((float)i + 0.5f) / (float)(ow * 2)));
      e->wsynxl[i] = e->wanxl[i] * e->wanxl[i] * e->wanxl[i];
      e->wsynxr[i] = e->wanxr[i] * e->wanxr[i] * e->wanxr[i];
    }
    for (int i = 0; i < oh; i++) {
      e->wanyl[i] = (float)sqrt((double)cosf(pi * ((float)(i - oh) + 0.5f) / (float)(oh * 2)));
      e->wanyr[i] = (float)sqrt((double)cosf(pi * ((float)i + 0.5f) / (float)(oh * 2)));
      e->wsynyl[i] = e->wanyl[i] * e->wanyl[i] * e->wanyl[i];
      e->wsynyr[i] = e->wanyr[i] * e->wanyr[i] * e->wanyr[i];
    }
  } else {
    for (int i = 0; i < ow; i++) {
      e->wanxl[i] = 1.0f; e->wanxr[i] = 1.0f;
      float a = cosf(pi * ((float)(i - ow) + 0.5f) / (float)(ow * 2));
      e->wsynxl[i] = a * a;
      float b = cosf(pi * ((float)i + 0.5f) / (float)(ow * 2));
      e->wsynxr[i] = b * b;
    }
    for (int i = 0; i < oh; i++) {
      e->wanyl[i] = 1.0f; e->wanyr[i] = 1.0f;
      float a = cosf(pi * ((float)(i - oh) + 0.5f) / (float)(oh * 2));
      e->wsynyl[i] = a * a;
      float b = cosf(pi * ((float)i + 0.5f) / (float)(oh * 2));
      e->wsynyr[i] = b * b;
    }
  }
}

/* Tables of src/fft3d_engine.h:285-361 and src/helper.cpp:14-43. */
static void make_tables(oengine* e) {
  int bw = e->ep.bw, bh = e->ep.bh, ow = e->outwidth;
  e->wsharpen = xcalloc(bh * ow, 4);
  e->wdehalo = xcalloc(bh * ow, 4);
  e->pwin = xcalloc(bh * ow, 4);
  for (int j = 0; j < bh; j++) {
    int dj = j >= bh / 2 ? bh - j : j;
    float d2v = (float)(dj * dj) * (e->ep.svr * e->ep.svr) / (float)((bh / 2) * (bh / 2));
    for (int i = 0; i < ow; i++) {
      float d2 = d2v + (float)(i * i) / (float)((bw / 2) * (bw / 2));
      e->wsharpen[j * ow + i] = (float)(1 - exp((double)(-d2 / (2 * e->ep.scutoff * e->ep.scutoff))));
    }
  }
  float wmax = 0;
  for (int j = 0; j < bh; j++) {
    int dj = j >= bh / 2 ? bh - j : j;
    float d2v = (float)(dj * dj) * (e->ep.svr * e->ep.svr) / (float)((bh / 2) * (bh / 2));
    for (int i = 0; i < ow; i++) {
      float d2 = d2v + (float)(i * i) / (float)((bw / 2) * (bw / 2));
      float v = (float)(exp((double)(-0.7f * d2 * e->ep.hr * e->ep.hr)) - exp((double)(-d2 * e->ep.hr * e->ep.hr)));
      e->wdehalo[j * ow + i] = v;
      if (v > wmax) wmax = v;
    }
  }
  for (int k = 0; k < bh * ow; k++) e->wdehalo[k] /= wmax;
  for (int j = 0; j < bh; j++) {
    int dj = j < bh - j ? j : bh - j;
    float fh = (float)dj * 2.0f / (float)bh;
    float fh2 = fh * fh;
    for (int i = 0; i < ow; i++) {
      float fw = (float)i * 2.0f / (float)bw;
      float fw2 = fw * fw;
      e->pwin[j * ow + i] = (fh2 + fw2) / (fh2 + fw2 + e->ep.pcutoff * e->ep.pcutoff);
    }
  }
  e->pattern2d = xcalloc(bh * ow, 8);
  e->pattern3d = xcalloc(bh * ow, 8);
  if ((e->ep.sigma2 != e->ep.sigma || e->ep.sigma3 != e->ep.sigma || e->ep.sigma4 != e->ep.sigma) && e->ep.pfactor == 0) {
    float ft2 = (float)sqrt((double)0.5f) / 2;
    float ft3 = (float)sqrt((double)0.5f) / 4;
    for (int h = 0; h < bh; h++)
      for (int w = 0; w < ow; w++) {
        float fy = ((float)bh - 2.0f * (float)abs(h - bh / 2)) / (float)bh;
        float fx = ((float)w * 1.0f) / (float)ow;
        float f = (float)sqrt((double)((fx * fx + fy * fy) * 0.5f));
        float s;
        if (f < ft3) s = e->ep.sigma4 + (e->ep.sigma3 - e->ep.sigma4) * f / ft3;
        else if (f < ft2) s = e->ep.sigma3 + (e->ep.sigma2 - e->ep.sigma3) * (f - ft3) / (ft2 - ft3);
        else s = e->ep.sigma + (e->ep.sigma2 - e->ep.sigma) * (1 - f) / (1 - ft2);
        e->pattern2d[h * ow + w] = (double)(s * s / e->norm);
      }
    e->pattern_set = 1;
    e->ep.pfactor = 1;
  }
}

/* Window weight of block element (i, j): product of the per-axis analysis
 * windows (src/buffer.cpp:137-378 applies the same weights region by region). */
static double wx_an(const oengine* e, int j) {
  int bw = e->ep.bw, ow = e->ep.ow;
  if (j < ow) return e->wanxl[j];
  if (j >= bw - ow) return e->wanxr[j - (bw - ow)];
  return 1.0;
}
static double wy_an(const oengine* e, int i) {
  int bh = e->ep.bh, oh = e->ep.oh;
  if (i < oh) return e->wanyl[i];
  if (i >= bh - oh) return e->wanyr[i - (bh - oh)];
  return 1.0;
}

static int reflect(int v, int n) {
  if (v < 0) v = -v;
  if (v >= n) v = 2 * n - 2 - v;
  return v;
}

/* Cover row -> source row of the processed region (src/buffer.cpp:5-80). */
static int cover_row_to_src(const oengine* e, int cy) {
  int ry = reflect(cy - e->mirh, e->Hc);
  if (!e->ep.interlaced) return ry;
  int half = e->Hc / 2;
  return ry < half ? 2 * ry : e->Hc - 1 - 2 * (ry - half);
}

static double read_px(const oengine* e, const uint8_t* base, int stride, int y, int x) {
  const uint8_t* row = base + (size_t)(y + e->tc) * stride;
  int xx = x + e->lc;
  if (e->fmt.is_float) return (double)((const float*)row)[xx];
  if (e->fmt.bytes == 1) return (double)row[xx];
  return (double)((const uint16_t*)row)[xx];
}

/* Forward spectra of all blocks of one plane: [nb][bh][outwidth]. */
static void plane_spectra(oengine* e, const uint8_t* src, int stride, cplx* out, int chroma_centre) {
  int bw = e->ep.bw, bh = e->ep.bh, ow = e->outwidth;
  int stepx = bw - e->ep.ow, stepy = bh - e->ep.oh;
  double base = (!e->fmt.is_float && chroma_centre) ? (double)(1 << (e->fmt.bits - 1)) : 0.0;
  for (int by = 0; by < e->noy; by++)
    for (int bx = 0; bx < e->nox; bx++) {
      for (int i = 0; i < bh; i++) {
        int sy = cover_row_to_src(e, by * stepy + i);
        double wy = wy_an(e, i);
        for (int j = 0; j < bw; j++) {
          int sx = reflect(bx * stepx + j - e->mirw, e->Wc);
          e->blk[i * bw + j] = (read_px(e, src, stride, sy, sx) - base) * wx_an(e, j) * wy;
        }
      }
      r2c2d(e->blk, out + (size_t)(by * e->nox + bx) * bh * ow, bw, bh, e->tmp);
    }
}

static void synth_overlap(oengine* e, const double* blocks, uint8_t* dst, int dstride, int chroma_centre) {
  int bw = e->ep.bw, bh = e->ep.bh, ow = e->ep.ow, oh = e->ep.oh;
  int stepx = bw - ow, stepy = bh - oh;
  double norm = e->norm;
  int is_float = e->fmt.is_float;
  double base = (!is_float && chroma_centre) ? (double)(1 << (e->fmt.bits - 1)) : 0.0;
  double maxv = is_float ? 1.0 : (double)((1 << e->fmt.bits) - 1);
  for (int y = 0; y < e->Hc; y++) {
    int ry;
    if (!e->ep.interlaced) ry = y;
    else ry = (y % 2 == 0) ? y / 2 : e->Hc / 2 + (e->Hc - 1 - y) / 2;
    int cy = ry + e->mirh;
    int r = cy / stepy, i = cy - r * stepy;
    uint8_t* orow = dst + (size_t)(y + e->tc) * dstride;
    for (int x = 0; x < e->Wc; x++) {
      int cx = x + e->mirw;
      int c = cx / stepx, j = cx - c * stepx;
      /* blocks: (r-1 or r, c-1 or c); element lookup */
#define B(bxx, byy, ii, jj) blocks[((size_t)((byy) * e->nox + (bxx)) * bh + (ii)) * bw + (jj)]
      double v;
      int yband = (i < oh) && r >= 1, xband = (j < ow) && c >= 1;
      if (yband && xband) {
        double wyr = (double)e->wsynyr[i] * norm, wyl = (double)e->wsynyl[i] * norm;
        double top = B(c - 1, r - 1, i + stepy, j + stepx) * e->wsynxr[j] + B(c, r - 1, i + stepy, j) * e->wsynxl[j];
        double bot = B(c - 1, r, i, j + stepx) * e->wsynxr[j] + B(c, r, i, j) * e->wsynxl[j];
        v = top * wyr + bot * wyl;
      } else if (yband) {
        double wyr = (double)e->wsynyr[i] * norm, wyl = (double)e->wsynyl[i] * norm;
        v = B(c, r - 1, i + stepy, j) * wyr + B(c, r, i, j) * wyl;
      } else if (xband) {
        v = (B(c - 1, r, i, j + stepx) * e->wsynxr[j] + B(c, r, i, j) * e->wsynxl[j]) * norm;
      } else {
        v = B(c, r, i, j) * norm;
      }
#undef B
      int xx = x + e->lc;
      if (is_float) {
        double o = v; if (o < 0) o = 0; if (o > 1) o = 1;
        ((float*)orow)[xx] = (float)o;
      } else {
        double t = v + 0.5 + base;
        long q = (long)t; /* truncation toward zero, as the reference's cast */
        if (q < 0) q = 0;
        if (q > (long)maxv) q = (long)maxv;
        if (e->fmt.bytes == 1) orow[xx] = (uint8_t)q; else ((uint16_t*)orow)[xx] = (uint16_t)q;
      }
    }
  }
}

/* -------------------------------------------------- spectral kernels -- */

static double wiener(double re, double im, double sig2, double lowlimit, double* fr, double* fi) {
  double psd = re * re + im * im + (double)1e-15f;
  double f = (psd - sig2) / psd;
  f = f > lowlimit ? f : lowlimit;
  *fr = re * f; *fi = im * f;
  return f;
}

static double s_factor(const oengine* e, double psd, int idx) {
  return 1 + (double)e->ep.sharpen * e->wsharpen[idx] *
             sqrt(psd * e->sigmaSquaredSharpenMaxNormed /
                  ((psd + e->sigmaSquaredSharpenMinNormed) * (psd + e->sigmaSquaredSharpenMaxNormed)));
}
static double d_factor(const oengine* e, double psd, int idx) {
  return (psd + e->ht2n) / ((psd + e->ht2n) + (double)e->ep.dehalo * e->wdehalo[idx] * psd);
}

/* src/code_impl/Apply.cpp:13-65 (+ code_impl_C.h:54-91 degrid fraction). */
static void apply2d(oengine* e, cplx* X, float sig2f, int pattern) {
  int bh = e->ep.bh, ow = e->outwidth;
  double lowlimit = (double)((e->ep.beta - 1) / e->ep.beta);
  for (int b = 0; b < e->nb; b++) {
    cplx* blk = X + (size_t)b * bh * ow;
    double gf = e->ep.degrid != 0 ? (double)e->ep.degrid * blk[0].re / e->grid[0].re : 0.0;
    for (int k = 0; k < bh * ow; k++) {
      double g0 = gf * e->grid[k].re, g1 = gf * e->grid[k].im;
      double cr = blk[k].re - g0, ci = blk[k].im - g1;
      double psd = cr * cr + ci * ci + (double)1e-15f;
      double sig2 = pattern ? (e->bpsd ? e->bpsd[(size_t)b * bh * ow + k] : e->pattern2d[k]) : (double)sig2f;
      double f = (psd - sig2) / psd;
      f = f > lowlimit ? f : lowlimit;
      if (!pattern) {
        if (e->ep.sharpen != 0 || e->ep.dehalo == 0) f *= s_factor(e, psd, k);
        if (e->ep.dehalo != 0) f *= d_factor(e, psd, k);
      }
      blk[k].re = cr * f + g0;
      blk[k].im = ci * f + g1;
    }
  }
}

/* src/code_impl/Sharpen.cpp:13-133. */
static void sharpen_pass(oengine* e, cplx* X) {
  if (e->ep.sharpen == 0 && e->ep.dehalo == 0) return;
  int bh = e->ep.bh, ow = e->outwidth;
  for (int b = 0; b < e->nb; b++) {
    cplx* blk = X + (size_t)b * bh * ow;
    double gf = e->ep.degrid != 0 ? (double)e->ep.degrid * blk[0].re / e->grid[0].re : 0.0;
    for (int k = 0; k < bh * ow; k++) {
      double g0 = gf * e->grid[k].re, g1 = gf * e->grid[k].im;
      double cr = blk[k].re - g0, ci = blk[k].im - g1;
      double psd = cr * cr + ci * ci + (double)1e-15f;
      double f = 1.0;
      if (e->ep.sharpen != 0) f *= s_factor(e, psd, k);
      if (e->ep.dehalo != 0) f *= d_factor(e, psd, k);
      blk[k].re = cr * f + g0;
      blk[k].im = ci * f + g1;
    }
  }
}

/* src/code_impl/Apply.cpp:67-235: length-bt DFT along t over frames
 * n+from..n+to, Wiener per bin, degrid x bt on the DC bin, output = inverse
 * DFT at the current frame = mean of the filtered bins. */
static void apply3d(oengine* e, cplx* const* in, int btc, int from, cplx* out, float sig2f, int pattern) {
  int bh = e->ep.bh, ow = e->outwidth;
  double lowlimit = (double)((e->ep.beta - 1) / e->ep.beta);
  cplx tw[5][5];
  for (int k = 0; k < btc; k++)
    for (int t = 0; t < btc; t++) {
      int tt = from + t;
      double a = -2.0 * M_PI * (double)(((long)k * tt % btc + btc) % btc) / btc;
      tw[k][t].re = cos(a); tw[k][t].im = sin(a);
    }
  int cur = -from;
  for (int b = 0; b < e->nb; b++) {
    size_t off = (size_t)b * bh * ow;
    double gf = e->ep.degrid != 0 ? (double)e->ep.degrid * in[cur][off].re / e->grid[0].re : 0.0;
    for (int k = 0; k < bh * ow; k++) {
      double g0 = gf * e->grid[k].re * btc, g1 = gf * e->grid[k].im * btc;
      double sr = 0, si = 0;
      for (int q = 0; q < btc; q++) {
        double fr = 0, fi = 0;
        for (int t = 0; t < btc; t++) {
          cplx v = in[t][off + k];
          fr += v.re * tw[q][t].re - v.im * tw[q][t].im;
          fi += v.re * tw[q][t].im + v.im * tw[q][t].re;
        }
        if (q == 0) { fr -= g0; fi -= g1; }
        double sig2 = pattern ? (e->bpsd ? e->bpsd[off + k] * (double)(float)btc : e->pattern3d[k]) : (double)sig2f;
        double wr, wi;
        wiener(fr, fi, sig2, lowlimit, &wr, &wi);
        sr += wr; si += wi;
      }
      out[off + k].re = (sr + g0) / btc;
      out[off + k].im = (si + g1) / btc;
    }
  }
}

/* src/code_impl/Kalman.cpp:13-67, state updated in place. */
static void kalman(oengine* e, const cplx* cur, int pattern) {
  int bh = e->ep.bh, ow = e->outwidth;
  double k2 = (double)(e->ep.kratio * e->ep.kratio);
  for (int b = 0; b < e->nb; b++)
    for (int k = 0; k < bh * ow; k++) {
      size_t o = (size_t)b * bh * ow + k;
      double sigma;
      if (pattern) {
        sigma = e->bpsd ? e->bpsd[o] : e->pattern2d[k];
        if (sigma < (double)1e-15f) sigma = (double)1e-15f;
      }
      else sigma = (double)e->sigmaSquaredNoiseNormed2D;
      cplx c = cur[o], p = e->outLast[o];
      double dr = c.re - p.re, di = c.im - p.im;
      if (dr * dr > sigma * k2 || di * di > sigma * k2) {
        e->covar[o].re = sigma; e->covar[o].im = sigma;
        e->covarProcess[o].re = sigma; e->covarProcess[o].im = sigma;
        e->outLast[o] = c;
      } else {
        double sr = e->covar[o].re + e->covarProcess[o].re;
        double si = e->covar[o].im + e->covarProcess[o].im;
        double gr = sr / (sr + sigma), gi = si / (si + sigma);
        e->covarProcess[o].re = gr * gr * sigma;
        e->covarProcess[o].im = gi * gi * sigma;
        e->covar[o].re = (1 - gr) * sr;
        e->covar[o].im = (1 - gi) * si;
        e->outLast[o].re = gr * c.re + (1 - gr) * p.re;
        e->outLast[o].im = gi * c.im + (1 - gi) * p.im;
      }
    }
}

/* src/helper.cpp:46-132. */
static void find_pattern_block(oengine* e, const cplx* X, int* px, int* py) {
  int bh = e->ep.bh, ow = e->outwidth;
  double best = 1e15f;
  for (int by = 2; by < e->noy - 2; by++)
    for (int bx = 2; bx < e->nox - 2; bx++) {
      const cplx* blk = X + (size_t)(e->nox * by + bx) * bh * ow;
      double gcur = (double)e->ep.degrid * blk[0].re / e->grid[0].re;
      double s = 0;
      for (int k = 0; k < bh * ow; k++) {
        double c0 = blk[k].re - gcur * e->grid[k].re, c1 = blk[k].im - gcur * e->grid[k].im;
        s += (c0 * c0 + c1 * c1) * e->pwin[k];
      }
      if (s < best) { *px = bx; *py = by; best = s; }
    }
}

/* Per-region patterns (no reference counterpart; the generalisation of
 * FindPatternBlock / SetPattern, src/helper.cpp:46-132): regions of
 * preg_x x preg_y blocks each take the pattern |X - g G|^2 pwin of their own
 * first strict minimum of the search sums over the blocks of the reference's
 * search range inside the region, else of the plane's choice (gx, gy); every
 * block of a region gets that pattern, as a per-block PSD. */
static double block_sum(const oengine* e, const cplx* X, int bx, int by) {
  int bh = e->ep.bh, ow = e->outwidth;
  const cplx* blk = X + (size_t)(e->nox * by + bx) * bh * ow;
  double gcur = (double)e->ep.degrid * blk[0].re / e->grid[0].re;
  double s = 0;
  for (int k = 0; k < bh * ow; k++) {
    double c0 = blk[k].re - gcur * e->grid[k].re, c1 = blk[k].im - gcur * e->grid[k].im;
    s += (c0 * c0 + c1 * c1) * e->pwin[k];
  }
  return s;
}

static void set_region_patterns(oengine* e, const cplx* X, int gx, int gy) {
  int bh = e->ep.bh, ow = e->outwidth, rx = e->preg_x, ry = e->preg_y;
  size_t n = (size_t)bh * ow;
  if (!e->bpsd) e->bpsd = xcalloc((size_t)e->nb * n, sizeof(double));
  for (int y0 = 0; y0 < e->noy; y0 += ry)
    for (int x0 = 0; x0 < e->nox; x0 += rx) {
      int x1 = x0 + rx < e->nox ? x0 + rx : e->nox, y1 = y0 + ry < e->noy ? y0 + ry : e->noy;
      int px = -1, py = -1;
      double best = 1e15f;
      for (int by = y0 > 2 ? y0 : 2; by < y1 && by < e->noy - 2; by++)
        for (int bx = x0 > 2 ? x0 : 2; bx < x1 && bx < e->nox - 2; bx++) {
          double v = block_sum(e, X, bx, by);
          if (v < best) { best = v; px = bx; py = by; }
        }
      if (px < 0) { px = gx; py = gy; }
      const cplx* blk = X + (size_t)(e->nox * py + px) * n;
      double gcur = (double)e->ep.degrid * blk[0].re / e->grid[0].re;
      for (size_t k = 0; k < n; k++) {
        double c0 = blk[k].re - gcur * e->grid[k].re, c1 = blk[k].im - gcur * e->grid[k].im;
        double v = (c0 * c0 + c1 * c1) * e->pwin[k];
        for (int by = y0; by < y1; by++)
          for (int bx = x0; bx < x1; bx++) e->bpsd[((size_t)by * e->nox + bx) * n + k] = v;
      }
    }
}

static void set_pattern(oengine* e, const cplx* X, int px, int py) {
  int bh = e->ep.bh, ow = e->outwidth;
  const cplx* blk = X + (size_t)(e->nox * py + px) * bh * ow;
  double gcur = (double)e->ep.degrid * blk[0].re / e->grid[0].re;
  for (int k = 0; k < bh * ow; k++) {
    double c0 = blk[k].re - gcur * e->grid[k].re, c1 = blk[k].im - gcur * e->grid[k].im;
    e->pattern2d[k] = (c0 * c0 + c1 * c1) * e->pwin[k];
  }
}

/* ------------------------------------------------------------ engine --- */

static const char* engine_init(oengine* e, const oparams* p, const oformat* f, int plane, int width, int height, int frames) {
  memset(e, 0, sizeof(*e));
  e->ep = *p;
  e->fmt = *f;
  e->frames = frames;
  e->preg_x = p->preg_x > 0 && p->preg_y > 0 ? p->preg_x : 0;
  e->preg_y = p->preg_x > 0 && p->preg_y > 0 ? p->preg_y : 0;
  e->chroma = f->family == 1 && plane != 0;
  e->framewidth = e->chroma ? width >> f->ssw : width;
  e->frameheight = e->chroma ? height >> f->ssh : height;
  float factor = f->bytes == 1 ? 1.0f : f->bytes == 2 ? (float)(1 << (f->bits - 8)) : 1 / 255.0f;
  e->ep.sigma *= factor; e->ep.sigma2 *= factor; e->ep.sigma3 *= factor; e->ep.sigma4 *= factor;
  e->ep.smin *= factor; e->ep.smax *= factor;
  if (e->ep.ow * 2 > e->ep.bw) return "ow must be less than bw / 2";
  if (e->ep.oh * 2 > e->ep.bh) return "oh must be less than bh / 2";
  if (e->ep.ow < 0) e->ep.ow = e->ep.bw / 3;
  if (e->ep.oh < 0) e->ep.oh = e->ep.bh / 3;
  if (e->ep.bt < -1 || e->ep.bt > 5) return "bt must be -1 (Sharpen), 0 (Kalman), 1..5 (Wiener)";
  int xs = e->chroma ? f->ssw : 0, ys = e->chroma ? f->ssh : 0;
  int bw = e->ep.bw, bh = e->ep.bh, ow = e->ep.ow, oh = e->ep.oh;
  e->nox = (((width - e->ep.l - e->ep.r) >> xs) - ow + (bw - ow - 1)) / (bw - ow) + 2;
  e->noy = (((height - e->ep.t - e->ep.b) >> ys) - oh + (bh - oh - 1)) / (bh - oh) + 2;
  e->mirw = bw - ow; e->mirh = bh - oh;
  if (e->ep.beta < 1) return "beta must be not less 1.0";
  e->coverwidth = e->nox * (bw - ow) + ow;
  e->coverheight = e->noy * (bh - oh) + oh;
  e->nb = e->nox * e->noy;
  e->outwidth = bw / 2 + 1;
  e->lc = e->chroma ? e->ep.l >> f->ssw : e->ep.l;
  e->rc = e->chroma ? e->ep.r >> f->ssw : e->ep.r;
  e->tc = e->chroma ? e->ep.t >> f->ssh : e->ep.t;
  e->bc = e->chroma ? e->ep.b >> f->ssh : e->ep.b;
  e->Wc = e->framewidth - e->lc - e->rc;
  e->Hc = e->frameheight - e->tc - e->bc;
  make_windows(e);
  e->norm = 1.0f / (float)(bw * bh);
  e->sigmaSquaredNoiseNormed2D = e->ep.sigma * e->ep.sigma / e->norm;
  e->sigmaSquaredSharpenMinNormed = e->ep.smin * e->ep.smin / e->norm;
  e->sigmaSquaredSharpenMaxNormed = e->ep.smax * e->ep.smax / e->norm;
  e->ht2n = e->ep.ht * e->ep.ht / e->norm;
  make_tables(e);
  size_t ssz = (size_t)e->nb * bh * e->outwidth;
  if (e->ep.bt == 0) {
    e->outLast = xcalloc(ssz, sizeof(cplx));
    e->covar = xcalloc(ssz, sizeof(cplx));
    e->covarProcess = xcalloc(ssz, sizeof(cplx));
    for (size_t k = 0; k < ssz; k++) {
      e->covar[k].re = e->covar[k].im = e->sigmaSquaredNoiseNormed2D;
      e->covarProcess[k].re = e->covarProcess[k].im = e->sigmaSquaredNoiseNormed2D;
    }
  }
  e->blk = xcalloc((size_t)bw * bh, 8);
  e->tmp = xcalloc((size_t)4 * (bw + bh) + 2 * (size_t)bh * e->outwidth + 16, sizeof(cplx));
  e->spec = xcalloc(frames, sizeof(cplx*));
  /* gridsample: spectrum of block (0,0) of a constant maximum-value cover,
   * without chroma centring (src/fft3d_engine.h:384-404). */
  double maxv = f->is_float ? 1.0 : (double)((1 << f->bits) - 1);
  for (int i = 0; i < bh; i++)
    for (int j = 0; j < bw; j++) e->blk[i * bw + j] = maxv * wx_an(e, j) * wy_an(e, i);
  e->grid = xcalloc((size_t)bh * e->outwidth, sizeof(cplx));
  r2c2d(e->blk, e->grid, bw, bh, e->tmp);
  return NULL;
}

static void engine_free(oengine* e) {
  free(e->wanxl); free(e->wanxr); free(e->wanyl); free(e->wanyr);
  free(e->wsynxl); free(e->wsynxr); free(e->wsynyl); free(e->wsynyr);
  free(e->wsharpen); free(e->wdehalo); free(e->pwin); free(e->pattern2d); free(e->pattern3d); free(e->bpsd);
  free(e->grid); free(e->outLast); free(e->covar); free(e->covarProcess);
  if (e->spec) for (int i = 0; i < e->frames; i++) free(e->spec[i]);
  free(e->spec); free(e->blk); free(e->tmp);
}

typedef struct {
  oparams p;
  oformat f;
  int width, height, frames;
  const uint8_t** planes; /* frames * nplanes */
  int strides[4];
  int process[4];
  int crop;
  oengine eng[4];
  int has_eng[4];
  int engine_count, copy_count;
} oclip;

static const cplx* spectrum_of(oclip* c, int pl, int n) {
  oengine* e = &c->eng[pl];
  if (n < 0) n = 0;
  if (n >= c->frames) n = c->frames - 1;
  if (!e->spec[n]) {
    e->spec[n] = xcalloc((size_t)e->nb * e->ep.bh * e->outwidth, sizeof(cplx));
    plane_spectra(e, c->planes[(size_t)n * c->f.nplanes + pl], c->strides[pl], e->spec[n], e->chroma);
  }
  return e->spec[n];
}

static void engine_frame(oclip* c, int pl, int n, uint8_t* dst, int dstride) {
  oengine* e = &c->eng[pl];
  int bh = e->ep.bh, ow = e->outwidth, bw = e->ep.bw;
  size_t ssz = (size_t)e->nb * bh * ow;
  const uint8_t* src = c->planes[(size_t)n * c->f.nplanes + pl];
  int sstride = c->strides[pl];
  int pattern = e->ep.pfactor != 0 || e->bpsd;
  for (int k = 0; k < n - 2 && k < e->frames; k++)
    if (e->spec[k] && k != e->ep.pframe) { free(e->spec[k]); e->spec[k] = NULL; }
  if (pattern && !e->bpsd && !e->pattern_set && !e->ep.pshow) {
    /* src/fft3d_engine.h:485-504 */
    int pf = e->ep.pframe;
    const cplx* ps = spectrum_of(c, pl, pf);
    int px = e->ep.px, py = e->ep.py;
    if (px == 0 && py == 0) find_pattern_block(e, ps, &px, &py);
    e->ep.px = px; e->ep.py = py;
    set_pattern(e, ps, px, py);
    if (e->preg_x > 0) set_region_patterns(e, ps, px, py);
    e->pattern_set = 1;
  }
  if (pattern && !e->bpsd && e->ep.pshow) {
    /* Pattern display, src/fft3d_engine.h:505-565: the pattern block of
     * this frame (searched unless px/py are given) reconstructed alone with
     * analysis and synthesis windows set to 1 -- permanently, so calls after
     * the first also search with flat windows -- and the plane centred for
     * every non-RGB plane, luma included. */
    cplx* X = xcalloc(ssz, sizeof(cplx));
    double* blk = xcalloc((size_t)e->nb * bh * bw, 8);
    plane_spectra(e, src, sstride, X, e->chroma);
    int px = e->ep.px, py = e->ep.py;
    if (px == 0 && py == 0) find_pattern_block(e, X, &px, &py);
    set_pattern(e, X, px, py);
    for (int i = 0; i < e->ep.ow; i++) e->wanxl[i] = e->wanxr[i] = e->wsynxl[i] = e->wsynxr[i] = 1.0f;
    for (int i = 0; i < e->ep.oh; i++) e->wanyl[i] = e->wanyr[i] = e->wsynyl[i] = e->wsynyr[i] = 1.0f;
    e->pshow_flat = 1;
    int centre = e->fmt.family != 2;
    plane_spectra(e, src, sstride, X, centre);
    int keep = py * e->nox + px;
    for (int b = 0; b < e->nb; b++)
      if (b != keep) memset(X + (size_t)b * bh * ow, 0, (size_t)bh * ow * sizeof(cplx));
    for (int b = 0; b < e->nb; b++) c2r2d(X + (size_t)b * bh * ow, blk + (size_t)b * bh * bw, bw, bh, e->tmp);
    synth_overlap(e, blk, dst, dstride, centre);
    free(X); free(blk);
    return;
  }
  cplx* out = xcalloc(ssz, sizeof(cplx));
  double* blocks = xcalloc((size_t)e->nb * bh * bw, 8);
  int btcur = e->ep.bt;
  if ((e->ep.bt / 2 > n) || (e->ep.bt - 1) / 2 > (e->frames - 1 - n)) btcur = 1;
  float sig2f = (float)btcur * e->ep.sigma * e->ep.sigma / e->norm;
  if (btcur > 0) {
    if (btcur == 1) {
      memcpy(out, spectrum_of(c, pl, n), ssz * sizeof(cplx));
      apply2d(e, out, sig2f, pattern);
      if (pattern) sharpen_pass(e, out);
    } else {
      if (!e->pattern3d_init && !e->bpsd) {
        for (int k = 0; k < bh * ow; k++) e->pattern3d[k] = e->pattern2d[k] * (double)(float)btcur;
        e->pattern3d_init = 1;
      }
      int from = -btcur / 2, to = (btcur - 1) / 2;
      cplx* in[5];
      for (int i = from; i <= to; i++) in[i - from] = (cplx*)spectrum_of(c, pl, n + i);
      apply3d(e, in, btcur, from, out, sig2f, pattern);
      sharpen_pass(e, out);
    }
  } else if (e->ep.bt == 0) {
    if (n == 0) {
      /* first frame passes through unfiltered (src/fft3d_engine.h:681-687) */
      for (int y = 0; y < e->frameheight; y++)
        memcpy(dst + (size_t)y * dstride, src + (size_t)y * sstride, (size_t)e->framewidth * e->fmt.bytes);
      free(out); free(blocks);
      return;
    }
    kalman(e, spectrum_of(c, pl, n), pattern);
    memcpy(out, e->outLast, ssz * sizeof(cplx));
    sharpen_pass(e, out);
  } else {
    memcpy(out, spectrum_of(c, pl, n), ssz * sizeof(cplx));
    sharpen_pass(e, out);
  }
  for (int b = 0; b < e->nb; b++) c2r2d(out + (size_t)b * bh * ow, blocks + (size_t)b * bh * bw, bw, bh, e->tmp);
  synth_overlap(e, blocks, dst, dstride, e->chroma);
  free(out); free(blocks);
}

static void parse_args(const char* args, oparams* p, int* have_planes) {
  const char* s = args ? args : "";
  *have_planes = 0;
  while (*s) {
    const char* semi = strchr(s, ';');
    size_t len = semi ? (size_t)(semi - s) : strlen(s);
    char item[256];
    if (len >= sizeof(item)) len = sizeof(item) - 1;
    memcpy(item, s, len); item[len] = 0;
    char* eq = strchr(item, '=');
    if (eq) {
      *eq = 0;
      const char* k = item; const char* v = eq + 1;
#define F(name) if (!strcmp(k, #name)) p->name = (float)atof(v);
#define I(name) if (!strcmp(k, #name)) p->name = atoi(v);
      F(sigma) F(beta) I(bw) I(bh) I(bt) I(ow) I(oh) F(kratio) F(sharpen) F(scutoff) F(svr) F(smin) F(smax)
      I(measure) I(interlaced) I(wintype) I(pframe) I(px) I(py) I(pshow) F(pcutoff) F(pfactor)
      F(sigma2) F(sigma3) F(sigma4) F(degrid) F(dehalo) F(hr) F(ht) I(l) I(t) I(r) I(b) I(opt) I(ncpu) I(mt)
#undef F
#undef I
      if (!strcmp(k, "pregion")) {
        p->preg_x = atoi(v);
        const char* comma = strchr(v, ',');
        p->preg_y = comma ? atoi(comma + 1) : p->preg_x;
      }
      if (!strcmp(k, "planes")) {
        *have_planes = 1;
        p->nplanes_list = 0;
        const char* q = v;
        while (*q && p->nplanes_list < 4) {
          p->planes[p->nplanes_list++] = atoi(q);
          const char* comma = strchr(q, ',');
          if (!comma) break;
          q = comma + 1;
        }
      }
    }
    if (!semi) break;
    s = semi + 1;
  }
}

static void set_err(char* err, int len, const char* msg) {
  if (err && len > 0) snprintf(err, (size_t)len, "FFT3D: %s", msg);
}

void nf3d_oracle_free(void* h);

/* Same signature as the reference driver's nf3dref_create. */
void* nf3d_oracle_create(const char* args, int family, int sample_type, int bits, int ssw, int ssh,
                         int width, int height, int nframes, const void* const* planes,
                         const int* strides, char* err, int errlen) {
  oclip* c = xcalloc(1, sizeof(oclip));
  /* families 3 / 4: YUVA / RGBA, a fourth (alpha) plane at the luma size */
  int alpha = family == 3 || family == 4;
  if (family == 3) family = 1;
  if (family == 4) family = 2;
  c->f.family = family;
  c->f.is_float = sample_type != 0;
  c->f.bits = bits;
  c->f.bytes = bits <= 8 ? 1 : bits <= 16 ? 2 : 4;
  c->f.ssw = family == 1 ? ssw : 0;
  c->f.ssh = family == 1 ? ssh : 0;
  c->f.nplanes = family == 0 ? 1 : 3 + alpha;
  c->width = width; c->height = height; c->frames = nframes;
  c->planes = xcalloc((size_t)nframes * c->f.nplanes, sizeof(void*));
  memcpy(c->planes, planes, sizeof(void*) * (size_t)nframes * c->f.nplanes);
  for (int i = 0; i < c->f.nplanes; i++) c->strides[i] = strides[i];
  /* defaults, src/fft3d.hpp:77-106 */
  float sigma1 = 2.0f;
  oparams p;
  memset(&p, 0, sizeof(p));
  {
    oparams tmp; int hp;
    memset(&tmp, 0, sizeof(tmp)); tmp.sigma = sigma1;
    parse_args(args, &tmp, &hp);
    sigma1 = tmp.sigma;
  }
  p.sigma = sigma1; p.beta = 1.0f; p.bw = 32; p.bh = 32; p.bt = 3; p.ow = -1; p.oh = -1;
  p.kratio = 2.0f; p.sharpen = 0; p.scutoff = 0.3f; p.svr = 1.0f; p.smin = 4.0f; p.smax = 20.0f;
  p.measure = 1; p.interlaced = 0; p.wintype = 0; p.pframe = 0; p.px = 0; p.py = 0; p.pshow = 0;
  p.pcutoff = 0.1f; p.pfactor = 0; p.sigma2 = sigma1; p.sigma3 = sigma1; p.sigma4 = sigma1;
  p.degrid = 1.0f; p.dehalo = 0; p.hr = 2.0f; p.ht = 50.0f; p.l = p.t = p.r = p.b = 0; p.opt = 0;
  p.ncpu = 2; p.mt = 0;
  int have_planes;
  parse_args(args, &p, &have_planes);
  if (p.l < 0) p.l = 0;
  if (p.t < 0) p.t = 0;
  if (p.r < 0) p.r = 0;
  if (p.b < 0) p.b = 0;
  c->p = p;
  c->crop = p.l > 0 || p.r > 0 || p.t > 0 || p.b > 0;
  if (p.l + p.r >= width) { set_err(err, errlen, "Width cannot be cropped to zero or below"); nf3d_oracle_free(c); return NULL; }
  if (p.t + p.b >= height) { set_err(err, errlen, "Height cannot be cropped to zero or below"); nf3d_oracle_free(c); return NULL; }
  for (int i = 0; i < 4; i++) c->process[i] = 2;
  if (!have_planes) { p.nplanes_list = 3; p.planes[0] = 0; p.planes[1] = 1; p.planes[2] = 2; }
  for (int i = 0; i < p.nplanes_list; i++)
    if (p.planes[i] < c->f.nplanes && p.planes[i] >= 0) c->process[p.planes[i]] = 3;
  /* An alpha plane is filtered with the chroma geometry of YUV formats
   * (src/fft3d.hpp:186-191: IsChroma for every plane but the first), which
   * for subsampled YUVA covers only part of the plane while the whole plane
   * is copied out (src/fft3d.hpp:292-313): the reference leaves the rest
   * undefined, so that case is refused. */
  if (c->f.nplanes == 4 && c->process[3] == 3 && c->f.family == 1 && (c->f.ssw || c->f.ssh)) {
    set_err(err, errlen, "the alpha plane of a subsampled YUVA clip cannot be filtered");
    nf3d_oracle_free(c);
    return NULL;
  }
  for (int i = 0; i < c->f.nplanes; i++) {
    if (c->process[i] == 3) {
      const char* msg = engine_init(&c->eng[i], &p, &c->f, i, width, height, nframes);
      c->has_eng[i] = 1;
      if (msg) { set_err(err, errlen, msg); nf3d_oracle_free(c); return NULL; }
      if (c->eng[i].ep.interlaced && (c->eng[i].Hc % 2)) { set_err(err, errlen, "interlaced requires an even plane height"); nf3d_oracle_free(c); return NULL; }
      c->engine_count++;
    } else if (c->process[i] == 2) c->copy_count++;
  }
  return c;
}

static void copy_plane_rows(uint8_t* dst, int dstride, const uint8_t* src, int sstride, size_t rowbytes, int rows) {
  for (int y = 0; y < rows; y++) memcpy(dst + (size_t)y * dstride, src + (size_t)y * sstride, rowbytes);
}

/* src/fft3d.hpp:221-352: per-plane engine output placed into the processed
 * region; crop borders and unprocessed planes copied from the source. */
int nf3d_oracle_get_frame(void* h, int n, void* const* dst, const int* dst_strides, char* err, int errlen) {
  oclip* c = (oclip*)h;
  (void)err; (void)errlen;
  for (int pl = 0; pl < c->f.nplanes; pl++) {
    int chroma = c->f.family == 1 && pl > 0 && pl < 3;
    int pw = chroma ? c->width >> c->f.ssw : c->width;
    int ph = chroma ? c->height >> c->f.ssh : c->height;
    const uint8_t* src = c->planes[(size_t)n * c->f.nplanes + pl];
    uint8_t* d = (uint8_t*)dst[pl];
    if (c->process[pl] == 3) {
      /* source everywhere first (crop borders), then the processed region */
      copy_plane_rows(d, dst_strides[pl], src, c->strides[pl], (size_t)pw * c->f.bytes, ph);
      engine_frame(c, pl, n, d, dst_strides[pl]);
    } else if (c->process[pl] == 2) {
      copy_plane_rows(d, dst_strides[pl], src, c->strides[pl], (size_t)pw * c->f.bytes, ph);
    }
  }
  return 0;
}

void nf3d_oracle_free(void* h) {
  oclip* c = (oclip*)h;
  if (!c) return;
  for (int i = 0; i < 4; i++) if (c->has_eng[i]) engine_free(&c->eng[i]);
  free(c->planes);
  free(c);
}

/* Per-block noise PSD ([nb][bh*outwidth], float) used in place of sigma and
 * the pattern by this plane's following frames (the learned-PSD form of the
 * library, whose model supplies exp(log-PSD) per block; 3D gains use it times
 * btcur); NULL switches back.  No reference counterpart. */
int nf3d_oracle_set_block_psd(void* h, int plane, const float* psd) {
  oclip* c = (oclip*)h;
  if (!c->has_eng[plane]) return -1;
  oengine* e = &c->eng[plane];
  size_t n = (size_t)e->nb * e->ep.bh * e->outwidth;
  if (!psd) { free(e->bpsd); e->bpsd = NULL; return 0; }
  if (!e->bpsd) e->bpsd = xcalloc(n, sizeof(double));
  for (size_t i = 0; i < n; i++) e->bpsd[i] = (double)psd[i];
  return 0;
}

/* Per-region pattern estimation for this plane (rx x ry blocks; 0 = the
 * reference's single pattern), before its first frame.  No reference
 * counterpart. */
int nf3d_oracle_set_pattern_regions(void* h, int plane, int rx, int ry) {
  oclip* c = (oclip*)h;
  if (!c->has_eng[plane] || rx < 0 || ry < 0) return -1;
  oengine* e = &c->eng[plane];
  e->preg_x = rx > 0 && ry > 0 ? rx : 0;
  e->preg_y = rx > 0 && ry > 0 ? ry : 0;
  return 0;
}

/* The per-block PSD the plane filters with (set by nf3d_oracle_set_block_psd
 * or estimated per region), [nb][bh*outwidth]; -1 when there is none.  No
 * reference counterpart. */
int nf3d_oracle_block_psd_get(void* h, int plane, double* out) {
  oclip* c = (oclip*)h;
  if (!c->has_eng[plane]) return -1;
  oengine* e = &c->eng[plane];
  if (!e->bpsd) return -1;
  memcpy(out, e->bpsd, (size_t)e->nb * e->ep.bh * e->outwidth * sizeof(double));
  return 0;
}

/* Per-plane geometry for tests: nox, noy, outwidth, mirw, mirh. */
int nf3d_oracle_geometry(void* h, int plane, int* out5) {
  oclip* c = (oclip*)h;
  if (!c->has_eng[plane]) return -1;
  oengine* e = &c->eng[plane];
  out5[0] = e->nox; out5[1] = e->noy; out5[2] = e->outwidth; out5[3] = e->mirw; out5[4] = e->mirh;
  return 0;
}

/* The plane's noise pattern (pattern2d, [bh*outwidth]) and the block it was
 * taken from, for tests; -1 while no pattern is set.  No reference
 * counterpart (the reference keeps them private in FFT3DEngine). */
int nf3d_oracle_pattern(void* h, int plane, double* out, int* pxpy) {
  oclip* c = (oclip*)h;
  if (!c->has_eng[plane]) return -1;
  oengine* e = &c->eng[plane];
  if (!e->pattern_set) return -1;
  memcpy(out, e->pattern2d, (size_t)e->ep.bh * e->outwidth * sizeof(double));
  pxpy[0] = e->ep.px;
  pxpy[1] = e->ep.py;
  return 0;
}
