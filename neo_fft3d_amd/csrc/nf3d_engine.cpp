// Per-plane engine of the MI355X denoiser: the C ABI's nf3d_engine_*.
//
// Host side of what FFT3DEngine does (src/fft3d_engine.h:87-741): parameter
// scaling and validation, block geometry, the window / sharpen / dehalo /
// pattern tables and the gridsample spectrum (built once, on the host, in the
// reference's float arithmetic -- the DFT of gridsample in double), and the
// per-frame orchestration of the gfx950 kernels:
//
//   bt >= 2 : spectra of the frames in [n - bt/2, n + (bt-1)/2] are kept in an
//             HBM ring keyed by frame number (replacing the LRU frame cache of
//             src/cache.hpp).  In order, the newest frame of the window is the
//             only one not yet transformed: MODE_W3D computes its spectrum on
//             chip, stores it to the ring once, and filters frame n with the
//             older spectra read from the ring.  Frames missing for any other
//             reason (start of the clip, random access) are transformed first
//             by MODE_FWD.
//   bt == 1, edge frames of bt >= 2 : MODE_W2D, all on chip.
//   bt == 0 : MODE_KALMAN with the recursion state resident in HBM.
//   bt == -1: MODE_SHARPEN.
// Every filtering launch ends with k_seam finishing the overlap-add bands
// shared between workgroup tiles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/nf3d.h"
#include "nf3d_internal.h"
#include "nf3d_mfma.h"
#include "nf3d_mlp_batch.h"

using nf3d::PatternJob;
using nf3d::PlaneJob;
using nf3d::RunLaunch;
using nf3d::TileLaunch;

namespace {

struct Failure {
  std::string msg;
};

[[noreturn]] void fail(const std::string& m) { throw Failure{m}; }

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    // clear the runtime's last error, so a later launch check does not
    // report this failure again (a failed allocation stays "last" otherwise)
    (void)hipGetLastError();
    fail(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
  }
}

void copy_err(char* err, size_t len, const std::string& m) {
  if (err && len) {
    size_t n = std::min(len - 1, m.size());
    memcpy(err, m.data(), n);
    err[n] = 0;
  }
}

// Launch tuning of the frame-loop path (results do not depend on it, except
// with a learned PSD, which is predicted once per launch: table entries do
// not apply to such engines): frames per k_run launch and, for the modes
// without a temporal recursion (Wiener 2D, sharpen), frames per workgroup.  A process-wide table keyed by block
// shape, bit depth and bt, loaded by nf3d_tune_load or from the file named
// by NF3D_TUNE_TABLE; engines take their entry when created.
struct TuneEntry {
  int bw, bh, bits, bt, frames, fpw;
};
std::mutex g_tune_mu;
std::vector<TuneEntry> g_tune;
bool g_tune_env_done = false;

int parse_tune_file(const char* path, std::vector<TuneEntry>& out, std::string& msg) {
  FILE* f = fopen(path, "r");
  if (!f) {
    msg = std::string("cannot open tuning table ") + path;
    return -1;
  }
  char line[256];
  int ln = 0;
  out.clear();
  while (fgets(line, sizeof line, f)) {
    ln++;
    const char* q = line;
    while (*q == ' ' || *q == '\t') q++;
    if (*q == '#' || *q == '\n' || *q == 0) continue;
    TuneEntry t{};
    if (sscanf(q, "%d %d %d %d %d %d", &t.bw, &t.bh, &t.bits, &t.bt, &t.frames, &t.fpw) != 6 || t.frames < 1 ||
        t.frames > nf3d::RUN_MAX_FRAMES || t.fpw < 1 || t.fpw > nf3d::RUN_MAX_FRAMES) {
      fclose(f);
      msg = std::string("tuning table ") + path + ": bad line " + std::to_string(ln);
      return -1;
    }
    out.push_back(t);
  }
  fclose(f);
  return (int)out.size();
}

bool tune_lookup(int bw, int bh, int bits, int bt, TuneEntry* hit) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  if (!g_tune_env_done) {
    g_tune_env_done = true;
    const char* path = getenv("NF3D_TUNE_TABLE");
    std::string msg;
    if (path && *path && parse_tune_file(path, g_tune, msg) < 0) fprintf(stderr, "nf3d: %s (ignored)\n", msg.c_str());
  }
  for (const TuneEntry& t : g_tune)
    if (t.bw == bw && t.bh == bh && t.bits == bits && t.bt == bt) {
      *hit = t;
      return true;
    }
  return false;
}

template <typename T>
T* dev_alloc(size_t count) {
  void* p = nullptr;
  check(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc");
  return static_cast<T*>(p);
}

}  // namespace

struct nf3d_engine {
  nf3d_params ep{};  // scaled copy, as FFT3DEngine keeps it
  nf3d_format fmt{};
  int plane = 0, chroma = 0, frames = 0, device = 0;
  int framewidth = 0, frameheight = 0, lc = 0, tc = 0, rc = 0, bc = 0, Wc = 0, Hc = 0;
  int nox = 0, noy = 0, nb = 0, outwidth = 0, mirw = 0, mirh = 0, pix = 1;
  float norm = 0, sigma2d = 0, smin_n = 0, smax_n = 0, ht2n = 0;
  std::vector<float> wanx, wany, wsynxl, wsynxr, wsynyl, wsynyr, wsharpen, wdehalo, pwin, pattern2d, pattern3d;
  std::vector<float2> grid;
  bool pattern_set = false, pattern3d_init = false;
  bool pshow_flat = false;  // pattern display has set the windows to 1
  float* d_ones = nullptr;  // flat window

  hipStream_t stream = nullptr;
  float* d_tab = nullptr;  // every float table, one allocation
  float *d_wanx, *d_wany, *d_wsynxl, *d_wsynxr, *d_wsynyl, *d_wsynyr, *d_wsharpen, *d_wdehalo, *d_pwin,
      *d_pattern2d, *d_pattern3d;
  float2* d_grid = nullptr;
  size_t spec_elems = 0;  // nb * bh * outwidth
  float2* d_ring = nullptr;
  int ring_slots = 0;
  std::vector<int> ring_frame;
  std::vector<char> ring_by_run;  // the slot's spectrum came from the frame-loop kernel
  std::vector<uint64_t> ring_tick;
  uint64_t tick = 0;
  float2 *d_kal_last = nullptr, *d_kal_covar = nullptr, *d_kal_cproc = nullptr;
  float* d_seam = nullptr;
  int ntx = 0;
  float2* d_spec_tmp = nullptr;
  float* d_psums = nullptr;
  int* d_choice = nullptr;
  // region_psd scratch (kept apart from the engine's own pattern estimate)
  float2* d_rspec = nullptr;
  float* d_rsums = nullptr;
  int* d_rchoice = nullptr;
  float* d_rglobal = nullptr;
  // staging for host-fed planes
  std::vector<uint8_t*> d_stage;
  std::vector<int> stage_frame;
  size_t stage_pitch = 0;
  int stage_next = 0;
  uint8_t* d_dst = nullptr;
  std::mutex mu;
  // Ordering between calls on different streams: the work of the last call
  // ends with done_ev on done_stream, and a call on another stream waits for
  // it first (the spectra ring, the Kalman state and the seam and staging
  // buffers carry over from one call to the next).
  hipEvent_t done_ev = nullptr;
  hipStream_t done_stream = nullptr;
  // opt-in HIP-event timing of the main spectral kernel of each frame
  bool profiling = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<int> ev_frames;  // frames filtered by each timed launch
  size_t ev_used = 0;
  // frame-loop launches (k_run): per-frame seam buffers, two launches' worth
  // (consecutive launches alternate, so a launch does not wait on the seam
  // pass of the one before for its buffer)
  float* d_seam_run = nullptr;
  static constexpr int kRunFrames = nf3d::RUN_MAX_FRAMES;  // largest k_run launch (frames)
  int run_frames = kRunFrames;  // frames per k_run launch (tuning)
  int run_fpw = 1;              // frames per workgroup, Wiener 2D / sharpen (tuning)
  bool run_frames_from_table = false;  // run_frames came from the tuning table, not the caller
  // 2D transforms: 0 = FFT in registers, 1 = DFT-matrix products on the matrix
  // cores (32x32 blocks in the frame-loop and the per-frame kernel, 64x64 in the
  // per-frame kernel)
  int dft_mode = 0;
  const void* dft_tab = nullptr;
  float dft_sx = 1.0f, dft_sz = 1.0f;
  // learned PSD: model, strength, the frame's spectra, features, per-block
  // pattern [nb][bh][outwidth] and the model's inference workspace
  nf3d_mlp* learned = nullptr;
  float learned_strength = 1.0f;
  // frame-loop predictions: 0 = one per launch (from its first frame); n > 0
  // = one per span of n frames [k n, (k + 1) n) (launches never cross a
  // span's start), from the span's first frame-loop frame of the call that
  // first reaches it; lspan = the span whose PSD d_lpat holds (-1: none), so
  // a call continuing a span reuses it
  int learned_period = 0;
  long long lspan = -1;
  // last frame of the latest device-range call (-1 after a per-frame call):
  // the frame shift a ring-bound graph replay is checked against
  int last_end = -1;
  float2* d_lspec = nullptr;
  void* d_lx = nullptr;
  float* d_lpat = nullptr;
  void* d_lwork = nullptr;
  // batched prediction for the engines of one frame-loop launch that share a
  // model (this engine first among them): features and PSDs of all their
  // blocks, rows in job order, and the model's workspace for that many rows
  void* d_bx = nullptr;
  float* d_bp = nullptr;
  void* d_bwork = nullptr;
  long long b_rows = 0;
  long long b_work = 0;
  std::vector<void*> b_retired;  // outgrown batch buffers (a captured graph may still use them)
  // per-region noise patterns (regions of preg_x x preg_y blocks; 0 = the
  // reference's single pattern): one pattern per block, 2D and 3D
  int preg_x = 0, preg_y = 0;
  float* d_rpat2d = nullptr;
  float* d_rpat3d = nullptr;

  ~nf3d_engine() {
    hipFree(d_tab);
    hipFree(d_grid);
    hipFree(d_ring);
    hipFree(d_kal_last);
    hipFree(d_kal_covar);
    hipFree(d_kal_cproc);
    hipFree(d_seam);
    hipFree(d_seam_run);
    hipFree(d_ones);
    hipFree(d_spec_tmp);
    hipFree(d_psums);
    hipFree(d_choice);
    hipFree(d_lspec);
    hipFree(d_lx);
    hipFree(d_lpat);
    hipFree(d_lwork);
    hipFree(d_bx);
    hipFree(d_bp);
    hipFree(d_bwork);
    for (void* q : b_retired) hipFree(q);
    hipFree(d_rpat2d);
    hipFree(d_rpat3d);
    hipFree(d_rspec);
    hipFree(d_rsums);
    hipFree(d_rchoice);
    hipFree(d_rglobal);
    for (auto* p : d_stage) hipFree(p);
    hipFree(d_dst);
    for (auto ev : ev_pool) hipEventDestroy(ev);
    if (done_ev) hipEventDestroy(done_ev);
    if (stream) hipStreamDestroy(stream);
  }

  // ------------------------------------------------------------- tables
  void build_windows() {
    // src/fft3d_engine.h:204-283, float arithmetic and glibc cosf
    const int ow = ep.ow, oh = ep.oh;
    const float pi = 3.1415926535897932384626433832795f;
    std::vector<float> axl(ow), axr(ow), ayl(oh), ayr(oh);
    wsynxl.assign(ow, 0);
    wsynxr.assign(ow, 0);
    wsynyl.assign(oh, 0);
    wsynyr.assign(oh, 0);
    auto cl = [&](int i, int o) { return cosf(pi * ((float)(i - o) + 0.5f) / (float)(o * 2)); };
    auto cr = [&](int i, int o) { return cosf(pi * ((float)i + 0.5f) / (float)(o * 2)); };
    if (ep.wintype == 0) {
      for (int i = 0; i < ow; i++) { axl[i] = cl(i, ow); axr[i] = cr(i, ow); wsynxl[i] = axl[i]; wsynxr[i] = axr[i]; }
      for (int i = 0; i < oh; i++) { ayl[i] = cl(i, oh); ayr[i] = cr(i, oh); wsynyl[i] = ayl[i]; wsynyr[i] = ayr[i]; }
    } else if (ep.wintype == 1) {
      for (int i = 0; i < ow; i++) {
        axl[i] = (float)std::sqrt((double)cl(i, ow));
        axr[i] = (float)std::sqrt((double)cr(i, ow));
        wsynxl[i] = axl[i] * axl[i] * axl[i];
        wsynxr[i] = axr[i] * axr[i] * axr[i];
      }
      for (int i = 0; i < oh; i++) {
        ayl[i] = (float)std::sqrt((double)cl(i, oh));
        ayr[i] = (float)std::sqrt((double)cr(i, oh));
        wsynyl[i] = ayl[i] * ayl[i] * ayl[i];
        wsynyr[i] = ayr[i] * ayr[i] * ayr[i];
      }
    } else {
      for (int i = 0; i < ow; i++) {
        axl[i] = axr[i] = 1.0f;
        float a = cl(i, ow), b = cr(i, ow);
        wsynxl[i] = a * a;
        wsynxr[i] = b * b;
      }
      for (int i = 0; i < oh; i++) {
        ayl[i] = ayr[i] = 1.0f;
        float a = cl(i, oh), b = cr(i, oh);
        wsynyl[i] = a * a;
        wsynyr[i] = b * b;
      }
    }
    wanx.assign(ep.bw, 1.0f);
    wany.assign(ep.bh, 1.0f);
    for (int j = 0; j < ow; j++) {
      wanx[j] = axl[j];
      wanx[ep.bw - ow + j] = axr[j];
    }
    for (int i = 0; i < oh; i++) {
      wany[i] = ayl[i];
      wany[ep.bh - oh + i] = ayr[i];
    }
  }

  void build_tables() {
    // src/fft3d_engine.h:285-361
    const int bw = ep.bw, bh = ep.bh, ow = outwidth;
    wsharpen.assign((size_t)bh * ow, 0);
    wdehalo.assign((size_t)bh * ow, 0);
    pwin.assign((size_t)bh * ow, 0);
    for (int j = 0; j < bh; j++) {
      int dj = j >= bh / 2 ? bh - j : j;
      float d2v = (float)(dj * dj) * (ep.svr * ep.svr) / (float)((bh / 2) * (bh / 2));
      for (int i = 0; i < ow; i++) {
        float d2 = d2v + (float)(i * i) / (float)((bw / 2) * (bw / 2));
        wsharpen[j * ow + i] = (float)(1 - std::exp((double)(-d2 / (2 * ep.scutoff * ep.scutoff))));
      }
    }
    float wmax = 0;
    for (int j = 0; j < bh; j++) {
      int dj = j >= bh / 2 ? bh - j : j;
      float d2v = (float)(dj * dj) * (ep.svr * ep.svr) / (float)((bh / 2) * (bh / 2));
      for (int i = 0; i < ow; i++) {
        float d2 = d2v + (float)(i * i) / (float)((bw / 2) * (bw / 2));
        float v = (float)(std::exp((double)(-0.7f * d2 * ep.hr * ep.hr)) - std::exp((double)(-d2 * ep.hr * ep.hr)));
        wdehalo[j * ow + i] = v;
        if (v > wmax) wmax = v;
      }
    }
    for (auto& v : wdehalo) v /= wmax;
    for (int j = 0; j < bh; j++) {
      int dj = std::min(j, bh - j);
      float fh = (float)dj * 2.0f / (float)bh;
      float fh2 = fh * fh;
      for (int i = 0; i < ow; i++) {
        float fw = (float)i * 2.0f / (float)bw;
        pwin[j * ow + i] = (fh2 + fw * fw) / (fh2 + fw * fw + ep.pcutoff * ep.pcutoff);
      }
    }
    pattern2d.assign((size_t)bh * ow, 0);
    pattern3d.assign((size_t)bh * ow, 0);
    if ((ep.sigma2 != ep.sigma || ep.sigma3 != ep.sigma || ep.sigma4 != ep.sigma) && ep.pfactor == 0) {
      // SigmasToPattern, src/helper.cpp:14-43
      float ft2 = (float)std::sqrt((double)0.5f) / 2;
      float ft3 = (float)std::sqrt((double)0.5f) / 4;
      for (int h = 0; h < bh; h++)
        for (int w = 0; w < ow; w++) {
          float fy = ((float)bh - 2.0f * (float)std::abs(h - bh / 2)) / (float)bh;
          float fx = ((float)w * 1.0f) / (float)ow;
          float f = (float)std::sqrt((double)((fx * fx + fy * fy) * 0.5f));
          float s;
          if (f < ft3) s = ep.sigma4 + (ep.sigma3 - ep.sigma4) * f / ft3;
          else if (f < ft2) s = ep.sigma3 + (ep.sigma2 - ep.sigma3) * (f - ft3) / (ft2 - ft3);
          else s = ep.sigma + (ep.sigma2 - ep.sigma) * (1 - f) / (1 - ft2);
          pattern2d[h * ow + w] = s * s / norm;
        }
      pattern_set = true;
      ep.pfactor = 1;
    }
  }

  // Spectrum of block (0,0) of a constant maximum-value cover without chroma
  // centring (src/fft3d_engine.h:384-404), with the gather's float products
  // (corner block: top rows (wx*wy)*p or (wy*p)*wx, band rows (wx*p)*wy) and a
  // double-precision DFT rounded once to float.
  void build_grid() {
    const int bw = ep.bw, bh = ep.bh, ow = ep.ow, oh = ep.oh;
    float maxv = fmt.is_float ? 1.0f : (float)((1 << fmt.bits_per_sample) - 1);
    std::vector<double> blk((size_t)bw * bh);
    for (int i = 0; i < bh; i++)
      for (int j = 0; j < bw; j++) {
        bool xw = j < ow || j >= bw - ow, yw = i < oh || i >= bh - oh;
        float wx = wanx[j], wy = wany[i], p = maxv, v;
        if (xw && yw) {
          bool yedge = i < oh;   // block row 0 (noy >= 2)
          bool xedge = j < ow;   // block column 0 (nox >= 2)
          if (yedge) v = xedge ? (wx * wy) * p : (wy * p) * wx;
          else v = (wx * p) * wy;
        } else if (xw) v = wx * p;
        else if (yw) v = wy * p;
        else v = p;
        blk[(size_t)i * bw + j] = v;
      }
    grid.assign((size_t)bh * outwidth, make_float2(0, 0));
    const double tp = 6.283185307179586476925286766559;
    for (int k1 = 0; k1 < bh; k1++)
      for (int k2 = 0; k2 < outwidth; k2++) {
        double sr = 0, si = 0;
        for (int i = 0; i < bh; i++)
          for (int j = 0; j < bw; j++) {
            double a = -tp * ((double)((long)k1 * i % bh) / bh + (double)((long)k2 * j % bw) / bw);
            sr += blk[(size_t)i * bw + j] * std::cos(a);
            si += blk[(size_t)i * bw + j] * std::sin(a);
          }
        grid[(size_t)k1 * outwidth + k2] = make_float2((float)sr, (float)si);
      }
  }

  void upload_tables() {
    const size_t sp = (size_t)ep.bh * outwidth;
    size_t total = ep.bw + ep.bh + 2 * ep.ow + 2 * ep.oh + 5 * sp + 16;
    d_tab = dev_alloc<float>(total);
    std::vector<float> h(total, 0.0f);
    size_t o = 0;
    auto put = [&](const std::vector<float>& v, float** dptr) {
      *dptr = d_tab + o;
      std::copy(v.begin(), v.end(), h.begin() + o);
      o += v.size();
    };
    put(wanx, &d_wanx);
    put(wany, &d_wany);
    put(wsynxl, &d_wsynxl);
    put(wsynxr, &d_wsynxr);
    put(wsynyl, &d_wsynyl);
    put(wsynyr, &d_wsynyr);
    put(wsharpen, &d_wsharpen);
    put(wdehalo, &d_wdehalo);
    put(pwin, &d_pwin);
    put(pattern2d, &d_pattern2d);
    put(pattern3d, &d_pattern3d);
    check(hipMemcpy(d_tab, h.data(), total * sizeof(float), hipMemcpyHostToDevice), "tables");
    d_grid = dev_alloc<float2>(sp);
    check(hipMemcpy(d_grid, grid.data(), sp * sizeof(float2), hipMemcpyHostToDevice), "grid");
  }

  // ------------------------------------------------------------ creation
  void init(const nf3d_params& p, const nf3d_format& f, int width, int height, int nframes, int pl) {
    ep = p;
    fmt = f;
    plane = pl;
    frames = nframes;
    chroma = f.color_family == 1 && pl != 0;
    framewidth = chroma ? width >> f.ssw : width;
    frameheight = chroma ? height >> f.ssh : height;
    pix = f.is_float ? 4 : f.bits_per_sample <= 8 ? 1 : 2;
    // src/fft3d_engine.h:91-136
    float factor = pix == 1 ? 1.0f : pix == 2 ? (float)(1 << (f.bits_per_sample - 8)) : 1 / 255.0f;
    ep.sigma *= factor;
    ep.sigma2 *= factor;
    ep.sigma3 *= factor;
    ep.sigma4 *= factor;
    ep.smin *= factor;
    ep.smax *= factor;
    if (ep.ow * 2 > ep.bw) fail("ow must be less than bw / 2");
    if (ep.oh * 2 > ep.bh) fail("oh must be less than bh / 2");
    if (ep.ow < 0) ep.ow = ep.bw / 3;
    if (ep.oh < 0) ep.oh = ep.bh / 3;
    if (ep.bt < -1 || ep.bt > 5) fail("bt must be -1 (Sharpen), 0 (Kalman), 1..5 (Wiener)");
    int xs = chroma ? f.ssw : 0, ys = chroma ? f.ssh : 0;
    const int bw = ep.bw, bh = ep.bh, ow = ep.ow, oh = ep.oh;
    if (bw <= 0 || bh <= 0 || bw - ow <= 0 || bh - oh <= 0) fail("block size must be positive");
    nox = (((width - ep.l - ep.r) >> xs) - ow + (bw - ow - 1)) / (bw - ow) + 2;
    noy = (((height - ep.t - ep.b) >> ys) - oh + (bh - oh - 1)) / (bh - oh) + 2;
    mirw = bw - ow;
    mirh = bh - oh;
    if (ep.beta < 1) fail("beta must be not less 1.0");
    nb = nox * noy;
    outwidth = bw / 2 + 1;
    lc = chroma ? ep.l >> f.ssw : ep.l;
    rc = chroma ? ep.r >> f.ssw : ep.r;
    tc = chroma ? ep.t >> f.ssh : ep.t;
    bc = chroma ? ep.b >> f.ssh : ep.b;
    Wc = framewidth - lc - rc;
    Hc = frameheight - tc - bc;
    // Cover padding is one reflect-101 of the plane (src/buffer.cpp:19-26,
    // 65-79); planes smaller than the padding would make the reference read
    // unfilled cover memory, so they are refused here.
    int coverw = nox * (bw - ow) + ow, coverh = noy * (bh - oh) + oh;
    if (Wc < 2 || Hc < 2 || mirw > Wc - 1 || mirh > Hc - 1 || coverw - mirw - Wc > Wc - 1 ||
        coverh - mirh - Hc > Hc - 1)
      fail("plane is too small for the block size");
    // every copied pixel must lie inside the block grid (an odd crop of a
    // subsampled plane can make the copied width exceed the geometry width)
    if ((mirw + Wc - 1) / (bw - ow) > nox - 1 || (mirh + Hc - 1) / (bh - oh) > noy - 1)
      fail("crop leaves the plane outside the block grid");
    if (ep.px < 0 || ep.py < 0 || ep.px >= nox || ep.py >= noy) fail("px/py outside the block grid");
    if (ep.interlaced && (Hc & 1)) fail("interlaced requires an even plane height");
    if (!nf3d::block_shape_supported(bw, bh)) fail("block size bw x bh not compiled for the GPU engine");
    build_windows();
    norm = 1.0f / (float)(bw * bh);
    sigma2d = ep.sigma * ep.sigma / norm;
    smin_n = ep.smin * ep.smin / norm;
    smax_n = ep.smax * ep.smax / norm;
    ht2n = ep.ht * ep.ht / norm;
    build_tables();
    build_grid();

    check(hipGetDevice(&device), "hipGetDevice");
    check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    upload_tables();
    spec_elems = (size_t)nb * bh * outwidth;
    d_seam = dev_alloc<float>((size_t)4 * Hc * Wc);
    int bpw = nf3d::tile_blocks_per_wg(bw, bh);
    ntx = (nox + bpw - 1) / bpw;
    TuneEntry tune;
    if (tune_lookup(bw, bh, fmt.is_float ? 32 : fmt.bits_per_sample, ep.bt, &tune)) {
      run_frames = tune.frames;
      run_fpw = tune.fpw;
      run_frames_from_table = true;
    }
    if (ep.bt >= 2) {
      ring_slots = ep.bt + 2;
      d_ring = dev_alloc<float2>(spec_elems * ring_slots);
      ring_frame.assign(ring_slots, -1);
      ring_by_run.assign(ring_slots, 0);
      ring_tick.assign(ring_slots, 0);
    }
    if (ep.bt == 0) {
      d_kal_last = dev_alloc<float2>(spec_elems);
      d_kal_covar = dev_alloc<float2>(spec_elems);
      d_kal_cproc = dev_alloc<float2>(spec_elems);
      std::vector<float2> h(spec_elems, make_float2(0, 0));
      check(hipMemcpy(d_kal_last, h.data(), spec_elems * sizeof(float2), hipMemcpyHostToDevice), "kalman");
      std::fill(h.begin(), h.end(), make_float2(sigma2d, sigma2d));
      check(hipMemcpy(d_kal_covar, h.data(), spec_elems * sizeof(float2), hipMemcpyHostToDevice), "kalman");
      check(hipMemcpy(d_kal_cproc, h.data(), spec_elems * sizeof(float2), hipMemcpyHostToDevice), "kalman");
    }
  }

  // ------------------------------------------------------------ per frame
  PlaneJob base_job() const {
    PlaneJob J{};
    J.nox = nox;
    J.noy = noy;
    J.nb = nb;
    J.ow = ep.ow;
    J.oh = ep.oh;
    J.mirw = mirw;
    J.mirh = mirh;
    J.Wc = Wc;
    J.Hc = Hc;
    J.interlaced = ep.interlaced;
    J.ntx = ntx;
    J.wg_begin = 0;
    J.seam_row_begin = 0;
    J.lc = lc;
    J.tc = tc;
    J.pix = pix;
    J.base = (chroma && !fmt.is_float) ? (float)(1 << (fmt.bits_per_sample - 1)) : 0.0f;
    J.obase = (chroma && !fmt.is_float) ? (1 << (fmt.bits_per_sample - 1)) : 0;
    J.maxval = fmt.is_float ? 1 : (1 << fmt.bits_per_sample) - 1;
    J.norm = norm;
    J.seam = d_seam;
    J.seam_plane = (long long)Hc * Wc;
    J.wanx = d_wanx;
    J.wany = d_wany;
    J.wsynxl = d_wsynxl;
    J.wsynxr = d_wsynxr;
    J.wsynyl = d_wsynyl;
    J.wsynyr = d_wsynyr;
    J.grid = d_grid;
    J.wsharpen = d_wsharpen;
    J.wdehalo = d_wdehalo;
    J.pattern2d = d_pattern2d;
    J.pattern3d = d_pattern3d;
    if (preg_x > 0) {  // per-region patterns, one per block
      J.pattern2d = d_rpat2d;
      J.pattern3d = d_rpat3d;
      J.pattern_per_block = 1;
    }
    J.sigma2d = sigma2d;
    J.lowlimit = (ep.beta - 1) / ep.beta;
    J.degrid = ep.degrid;
    J.sharpen = ep.sharpen;
    J.smin_n = smin_n;
    J.smax_n = smax_n;
    J.dehalo = ep.dehalo;
    J.ht2n = ht2n;
    J.kratio2 = ep.kratio * ep.kratio;
    J.use_pattern = ep.pfactor != 0;
    J.use_degrid = ep.degrid != 0;
    J.kal_last = d_kal_last;
    J.kal_covar = d_kal_covar;
    J.kal_cproc = d_kal_cproc;
    J.dft_tab = dft_mode ? dft_tab : nullptr;
    J.dft_sx = dft_sx;
    J.dft_sz = dft_sz;
    return J;
  }

  struct Src {
    const uint8_t* p;
    long long pitch;
  };

  struct Call {
    const nf3d_frame_ref* frames;
    int nframes;
    bool device;
    hipStream_t s;
  };

  const nf3d_frame_ref* find(const Call& c, int k) const {
    for (int i = 0; i < c.nframes; i++)
      if (c.frames[i].n == k) return &c.frames[i];
    return nullptr;
  }

  Src source(const Call& c, int k) {
    const nf3d_frame_ref* fr = find(c, k);
    if (!fr || !fr->data) fail("frame " + std::to_string(k) + " was not provided");
    if (c.device) return Src{static_cast<const uint8_t*>(fr->data), (long long)fr->stride};
    // host plane -> staging slot
    for (size_t i = 0; i < stage_frame.size(); i++)
      if (stage_frame[i] == k) return Src{d_stage[i], (long long)stage_pitch};
    if (d_stage.empty()) {
      stage_pitch = ((size_t)framewidth * pix + 255) / 256 * 256;
      for (int i = 0; i < 7; i++) d_stage.push_back(dev_alloc<uint8_t>(stage_pitch * frameheight));
      stage_frame.assign(d_stage.size(), -1);
    }
    int slot = stage_next;
    stage_next = (stage_next + 1) % (int)d_stage.size();
    stage_frame[slot] = k;
    check(hipMemcpy2DAsync(d_stage[slot], stage_pitch, fr->data, fr->stride, (size_t)framewidth * pix,
                           frameheight, hipMemcpyHostToDevice, c.s),
          "upload");
    return Src{d_stage[slot], (long long)stage_pitch};
  }

  // Start / end events of a timed launch (null when profiling is off).
  std::pair<hipEvent_t, hipEvent_t> timer(bool timed, int nframes, hipStream_t s) {
    if (!timed || !profiling) return {nullptr, nullptr};
    while (ev_pool.size() < ev_used + 2) {
      hipEvent_t ev;
      check(hipEventCreate(&ev), "hipEventCreate");
      ev_pool.push_back(ev);
    }
    ev_frames.resize(ev_pool.size() / 2);
    ev_frames[ev_used / 2] = nframes;
    hipEvent_t a = ev_pool[ev_used++], b = ev_pool[ev_used++];
    check(hipEventRecord(a, s), "hipEventRecord");
    return {a, b};
  }

  // A call on stream s starts after the previous call's work (if that went
  // to another stream) ...
  void order_before(hipStream_t s) {
    if (done_stream && done_stream != s) check(hipStreamWaitEvent(s, done_ev, 0), "hipStreamWaitEvent");
  }
  // ... and marks its own end.
  void order_after(hipStream_t s) {
    if (!done_ev) check(hipEventCreateWithFlags(&done_ev, hipEventDisableTiming), "hipEventCreate");
    check(hipEventRecord(done_ev, s), "hipEventRecord");
    done_stream = s;
  }

  void launch(int mode, int bt, const PlaneJob& J, hipStream_t s, bool timed = false, int onchip = 0) {
    TileLaunch tl{};
    tl.job[0] = J;
    tl.njobs = 1;
    tl.total_wg = noy * ntx;
    tl.total_seam_rows = Hc;
    tl.max_wc = Wc;
    auto ev = timer(timed, 1, s);
    check(nf3d::launch_tile(mode, bt, onchip, dft_mode, ep.bw, ep.bh, tl, s), "tile kernel");
    if (ev.second) check(hipEventRecord(ev.second, s), "hipEventRecord");
  }

  // Sum of the timed launches' durations (ms) and the frames they filtered
  // since the last read.
  double profile_read(int* nframes) {
    double total = 0;
    int n = 0;
    for (size_t i = 0; i + 1 < ev_used; i += 2) {
      check(hipEventSynchronize(ev_pool[i + 1]), "hipEventSynchronize");
      float ms = 0;
      check(hipEventElapsedTime(&ms, ev_pool[i], ev_pool[i + 1]), "hipEventElapsedTime");
      total += ms;
      n += ev_frames[i / 2];
    }
    if (nframes) *nframes = n;
    ev_used = 0;
    return total;
  }

  int ring_lookup(int k) {
    for (int i = 0; i < ring_slots; i++)
      if (ring_frame[i] == k) {
        ring_tick[i] = ++tick;
        return i;
      }
    return -1;
  }

  int ring_claim(int k, int lo, int hi) {
    int best = -1;
    for (int i = 0; i < ring_slots; i++) {
      if (ring_frame[i] >= lo && ring_frame[i] <= hi) continue;
      if (best < 0 || ring_tick[i] < ring_tick[best]) best = i;
    }
    ring_frame[best] = k;
    ring_by_run[best] = 0;
    ring_tick[best] = ++tick;
    return best;
  }

  float2* slot_ptr(int s) { return d_ring + spec_elems * (size_t)s; }

  void estimate_pattern(const Call& c) {
    // src/fft3d_engine.h:485-504 with FindPatternBlock / SetPattern
    int pf = std::max(0, std::min(ep.pframe, frames - 1));
    Src ps = source(c, pf);
    if (!d_spec_tmp) {
      d_spec_tmp = dev_alloc<float2>(spec_elems);
      d_psums = dev_alloc<float>(nb);
      d_choice = dev_alloc<int>(2);
    }
    PlaneJob J = base_job();
    J.src = ps.p;
    J.src_pitch = ps.pitch;
    J.spec_out = d_spec_tmp;
    launch(nf3d::MODE_FWD, 0, J, c.s);
    PatternJob P{};
    P.nox = nox;
    P.noy = noy;
    P.bh = ep.bh;
    P.outwidth = outwidth;
    P.spec = d_spec_tmp;
    P.grid = d_grid;
    P.pwin = d_pwin;
    P.degrid = ep.degrid;
    P.px = ep.px;
    P.py = ep.py;
    P.pattern2d = d_pattern2d;
    P.sums = d_psums;
    P.choice = d_choice;
    check(nf3d::launch_pattern(P, c.s), "pattern kernels");
    if (preg_x > 0) {
      // the 3D pattern is the 2D one times the first 3D btcur, which is bt
      // (src/fft3d_engine.h:625-630)
      check(nf3d::launch_pattern_regions(P, preg_x, preg_y, d_rpat2d, d_rpat3d, ep.bt >= 2 ? (float)ep.bt : 1.0f,
                                         c.s),
            "region pattern kernel");
    }
    int choice[2];
    check(hipMemcpyAsync(choice, d_choice, sizeof(choice), hipMemcpyDeviceToHost, c.s), "pattern");
    check(hipMemcpyAsync(pattern2d.data(), d_pattern2d, pattern2d.size() * sizeof(float),
                         hipMemcpyDeviceToHost, c.s),
          "pattern");
    check(hipStreamSynchronize(c.s), "pattern");
    ep.px = choice[0];
    ep.py = choice[1];
    pattern_set = true;
  }

  // The per-region noise patterns of one plane as a per-block PSD: the
  // plane's analysis, the plane-wide search (the fallback of regions without
  // a searchable block), then each region's own search and its block's
  // |X - degrid correction|^2 x pattern window written to every block of the
  // region (src/helper.cpp:46-132, region by region).  Leaves the engine's
  // own pattern untouched.
  void region_psd(const uint8_t* src, long long pitch, int rx, int ry, float* out, hipStream_t s) {
    if (!d_rspec) {
      d_rspec = dev_alloc<float2>(spec_elems);
      d_rsums = dev_alloc<float>(nb);
      d_rchoice = dev_alloc<int>(2);
      d_rglobal = dev_alloc<float>((size_t)ep.bh * outwidth);
    }
    PlaneJob J = base_job();
    J.src = src;
    J.src_pitch = pitch;
    J.spec_out = d_rspec;
    launch(nf3d::MODE_FWD, 0, J, s);
    PatternJob P{};
    P.nox = nox;
    P.noy = noy;
    P.bh = ep.bh;
    P.outwidth = outwidth;
    P.spec = d_rspec;
    P.grid = d_grid;
    P.pwin = d_pwin;
    P.degrid = ep.degrid;
    P.px = 0;
    P.py = 0;
    P.pattern2d = d_rglobal;
    P.sums = d_rsums;
    P.choice = d_rchoice;
    check(nf3d::launch_pattern(P, s), "pattern kernels");
    check(nf3d::launch_pattern_regions(P, rx, ry, out, out, 1.0f, s), "region pattern kernel");
  }

  void ensure_pattern3d(int btcur, hipStream_t s) {
    // Pattern2Dto3D with the first 3D btcur (src/fft3d_engine.h:625-630)
    if (pattern3d_init) return;
    for (size_t i = 0; i < pattern2d.size(); i++) pattern3d[i] = pattern2d[i] * (float)btcur;
    check(hipMemcpyAsync(d_pattern3d, pattern3d.data(), pattern3d.size() * sizeof(float),
                         hipMemcpyHostToDevice, s),
          "pattern3d");
    check(hipStreamSynchronize(s), "pattern3d");
    pattern3d_init = true;
  }

  // Pattern display, src/fft3d_engine.h:505-565: frame n's pattern block
  // (searched unless px/py are given) reconstructed alone, with analysis and
  // synthesis windows set to 1 -- permanently: calls after the first search
  // with flat windows too -- and the plane centred for every non-RGB plane,
  // luma included.
  void run_pshow(int n, const Call& c, uint8_t* dst, long long dst_pitch) {
    Src s0 = source(c, n);
    if (!d_spec_tmp) {
      d_spec_tmp = dev_alloc<float2>(spec_elems);
      d_psums = dev_alloc<float>(nb);
      d_choice = dev_alloc<int>(2);
    }
    if (!d_ones) {
      std::vector<float> ones(std::max(ep.bw, ep.bh), 1.0f);
      d_ones = dev_alloc<float>(ones.size());
      check(hipMemcpy(d_ones, ones.data(), ones.size() * sizeof(float), hipMemcpyHostToDevice), "ones");
    }
    auto flat = [&](PlaneJob& J) {
      J.wanx = J.wany = J.wsynxl = J.wsynxr = J.wsynyl = J.wsynyr = d_ones;
    };
    PlaneJob F = base_job();
    if (pshow_flat) flat(F);
    F.src = s0.p;
    F.src_pitch = s0.pitch;
    F.spec_out = d_spec_tmp;
    launch(nf3d::MODE_FWD, 0, F, c.s);
    PatternJob P{};
    P.nox = nox;
    P.noy = noy;
    P.bh = ep.bh;
    P.outwidth = outwidth;
    P.spec = d_spec_tmp;
    P.grid = d_grid;
    P.pwin = d_pwin;
    P.degrid = ep.degrid;
    P.px = ep.px;
    P.py = ep.py;
    P.pattern2d = d_pattern2d;
    P.sums = d_psums;
    P.choice = d_choice;
    check(nf3d::launch_pattern(P, c.s), "pattern kernels");
    pshow_flat = true;
    PlaneJob J = base_job();
    flat(J);
    const bool centre = fmt.color_family != 2 && !fmt.is_float;
    J.base = centre ? (float)(1 << (fmt.bits_per_sample - 1)) : 0.0f;
    J.obase = centre ? (1 << (fmt.bits_per_sample - 1)) : 0;
    J.sharpen = 0.0f;
    J.dehalo = 0.0f;
    J.keep_choice = d_choice;
    J.src = s0.p;
    J.src_pitch = s0.pitch;
    J.dst = dst;
    J.dst_pitch = dst_pitch;
    launch(nf3d::MODE_SHARPEN, 0, J, c.s, true);
  }

  // Learned PSD of one plane's blocks: the plane's own analysis, features,
  // exp(model log-PSD) * mult * strength into d_lpat.
  void learned_psd_plane(const uint8_t* src, long long pitch, float mult, hipStream_t s) {
    PlaneJob F = base_job();
    F.src = src;
    F.src_pitch = pitch;
    F.spec_out = d_lspec;
    launch(nf3d::MODE_FWD, 0, F, s);
    learned_psd_spec(d_lspec, mult, s);
  }

  // ... from spectra already computed (a frame in the Wiener-3D ring)
  void learned_psd_spec(const float2* spec, float mult, hipStream_t s) {
    lspan = -1;
    if (nf3d_mlp_features(learned, spec, nb, d_lx, s)) fail("learned PSD: feature kernel");
    char msg[256] = {0};
    if (nf3d_mlp_predict(learned, d_lx, nb, d_lpat, 1, mult * learned_strength, d_lwork, s, msg, sizeof msg))
      fail(std::string("learned PSD: ") + msg);
  }

  void use_learned(PlaneJob& J) const {
    J.use_pattern = 1;
    J.pattern_per_block = 1;
    J.pattern2d = d_lpat;
    J.pattern3d = d_lpat;
  }

  // Per-frame path: the PSD of frame n itself.
  void learned_psd(int n, const Call& c, float mult, PlaneJob& J) {
    Src s0 = source(c, n);
    learned_psd_plane(s0.p, s0.pitch, mult, c.s);
    use_learned(J);
  }

  void run(int n, const Call& c, uint8_t* dst, long long dst_pitch) {
    check(hipSetDevice(device), "hipSetDevice");
    if (ep.pfactor != 0 && ep.pshow) {
      run_pshow(n, c, dst, dst_pitch);
      return;
    }
    if (ep.pfactor != 0 && !pattern_set && !learned) estimate_pattern(c);
    int btcur = ep.bt;
    if ((ep.bt / 2 > n) || (ep.bt - 1) / 2 > (frames - 1 - n)) btcur = 1;
    PlaneJob J = base_job();
    J.sigma2 = (float)btcur * ep.sigma * ep.sigma / norm;
    J.dst = dst;
    J.dst_pitch = dst_pitch;
    if (ep.bt == 0 && n == 0) {
      // the first frame passes through unfiltered (src/fft3d_engine.h:681-687)
      Src s0 = source(c, 0);
      check(hipMemcpy2DAsync(dst + (long long)tc * dst_pitch + (long long)lc * pix, dst_pitch,
                             s0.p + (long long)tc * s0.pitch + (long long)lc * pix, s0.pitch,
                             (size_t)Wc * pix, Hc, hipMemcpyDeviceToDevice, c.s),
            "copy");
      return;
    }
    if (btcur > 0) {
      if (btcur == 1) {
        Src s0 = source(c, n);
        J.src = s0.p;
        J.src_pitch = s0.pitch;
        // keep the spectrum for the neighbouring 3D windows (results-neutral)
        if (ring_slots && ring_lookup(n) < 0) J.spec_out = slot_ptr(ring_claim(n, n, n));
        if (learned) learned_psd(n, c, 1.0f, J);
        launch(nf3d::MODE_W2D, 0, J, c.s, true);
      } else {
        if (J.use_pattern && !learned) ensure_pattern3d(btcur, c.s);
        if (learned) learned_psd(n, c, (float)btcur, J);
        const int from = -btcur / 2, to = (btcur - 1) / 2;
        for (int i = from; i < to; i++) {  // older frames: from the ring, else transformed now
          int k = n + i;
          int slot = ring_lookup(k);
          if (slot < 0) {
            slot = ring_claim(k, n + from, n + to);
            Src sk = source(c, k);
            PlaneJob F = base_job();
            F.src = sk.p;
            F.src_pitch = sk.pitch;
            F.spec_out = slot_ptr(slot);
            launch(nf3d::MODE_FWD, 0, F, c.s);
          }
          J.spec_in[i - from] = slot_ptr(slot);
        }
        int newest = ring_lookup(n + to);
        int onchip = newest < 0;
        if (onchip) {
          Src sk = source(c, n + to);
          J.src = sk.p;
          J.src_pitch = sk.pitch;
          J.spec_out = slot_ptr(ring_claim(n + to, n + from, n + to));
        } else {
          J.spec_in[btcur - 1] = slot_ptr(newest);
        }
        launch(nf3d::MODE_W3D, btcur, J, c.s, true, onchip);
      }
    } else {
      if (learned && ep.bt == 0) learned_psd(n, c, 1.0f, J);
      Src s0 = source(c, n);
      J.src = s0.p;
      J.src_pitch = s0.pitch;
      launch(ep.bt == 0 ? nf3d::MODE_KALMAN : nf3d::MODE_SHARPEN, 0, J, c.s, true);
    }
  }

  // One frame of a device-resident clip through the per-frame path, with
  // its temporal window (and the pattern frame while the pattern is pending).
  void run_one(int n, const void* const* src, long long src_pitch, void* const* dst, long long dst_pitch,
               hipStream_t s) {
    nf3d_frame_ref win[6];
    int k = 0;
    const int lo = ep.bt > 1 ? std::max(n - ep.bt / 2, 0) : n;
    const int hi = ep.bt > 1 ? std::min(n + (ep.bt - 1) / 2, frames - 1) : n;
    for (int i = lo; i <= hi; i++) win[k++] = nf3d_frame_ref{i, src[i], (int)src_pitch};
    if (ep.pfactor != 0 && !pattern_set && !learned) {
      const int pf = std::max(0, std::min(ep.pframe, frames - 1));
      if (pf < lo || pf > hi) win[k++] = nf3d_frame_ref{pf, src[pf], (int)src_pitch};
    }
    Call c{win, k, true, s};
    run(n, c, static_cast<uint8_t*>(dst[n]), dst_pitch);
  }

  int shp_kind() const {
    if (ep.sharpen == 0.0f && ep.dehalo == 0.0f) return nf3d::SHP_NONE;
    return ep.dehalo == 0.0f ? nf3d::SHP_SHARPEN : nf3d::SHP_GENERAL;
  }

  // Plan of a device range: frames [first, first + count) of a device-resident
  // clip, in order.  The frames whose filter window is whole (every frame for
  // bt = 1 / -1, all but the first for Kalman, btcur == bt for Wiener 3D),
  // frames [a, b], go through k_run in launches of up to kRunFrames frames
  // (`launches`, one job record each); the others (clip ends) through the
  // per-frame path, before (prologue) and after (epilogue) them.
  struct RangePlan {
    int first = 0, count = 0, a = 0, b = -1;
    int mode = 0, shp = 0, from = 0, to = 0;
    bool run = false;  // false: every frame through the per-frame path
    const void* const* src = nullptr;
    void* const* dst = nullptr;
    long long src_pitch = 0, dst_pitch = 0;
    std::vector<RunLaunch> launches;
    // learned PSD: launch j predicts (else it keeps its span's PSD in d_lpat)
    std::vector<char> predict;
    // Wiener 3D: the clip-end frames of the range (btcur 1, the 2D gains)
    // as one-frame launches of the frame-loop kernel in its 2D mode, instead
    // of the per-frame path: those before the frame-loop frames (run after
    // the prologue, while the ring still holds them) and those after
    std::vector<RunLaunch> edges_start, edges;
  };

  RangePlan plan_range(int first, int count, const void* const* src, long long src_pitch, void* const* dst,
                       long long dst_pitch) {
    for (int n = first; n < first + count; n++)
      if (n < 0 || n >= frames) fail("frame " + std::to_string(n) + " outside the clip");
    RangePlan P;
    P.first = first;
    P.count = count;
    P.src = src;
    P.dst = dst;
    P.src_pitch = src_pitch;
    P.dst_pitch = dst_pitch;
    int lo = 0, hi = frames - 1;
    P.mode = nf3d::MODE_W2D;
    if (ep.bt >= 2) {
      P.mode = nf3d::MODE_W3D;
      lo = ep.bt / 2;
      hi = frames - 1 - (ep.bt - 1) / 2;
    } else if (ep.bt == 0) {
      P.mode = nf3d::MODE_KALMAN;
      lo = 1;
    } else if (ep.bt == -1) {
      P.mode = nf3d::MODE_SHARPEN;
    }
    P.shp = shp_kind();
    P.a = std::max(first, lo);
    P.b = std::min(first + count - 1, hi);
    P.from = P.mode == nf3d::MODE_W3D ? -(ep.bt / 2) : 0;
    P.to = P.mode == nf3d::MODE_W3D ? (ep.bt - 1) / 2 : 0;
    // k_run keeps 32-bit offsets (source rows, destination and seam
    // elements); the sharpen-only mode without sharpen or dehalo is a plain
    // round trip, and the pattern display a different path: both are left
    // to the per-frame path
    const bool fits = (long long)(tc + Hc) * src_pitch < (1LL << 31) &&
                      (long long)(tc + Hc) * dst_pitch < (1LL << 31) && 4LL * Hc * Wc < (1LL << 31) &&
                      dst_pitch % pix == 0;
    const bool pshow = ep.pfactor != 0 && ep.pshow;
    // the matrix-core form of 64x64 blocks runs in the per-frame kernel only
    // (in the frame loop its operands would not fit the registers)
    const bool mf_tile = dft_mode && ep.bw == 64;
    P.run = P.a <= P.b && fits && !pshow && !mf_tile && !(P.mode == nf3d::MODE_SHARPEN && P.shp == nf3d::SHP_NONE);
    if (!P.run) return P;
    // launches of equal length (within one frame), two halves of seam
    // buffers in turn
    const size_t seam_frame = (size_t)4 * Hc * Wc;
    if (!d_seam_run) d_seam_run = dev_alloc<float>(2 * kRunFrames * seam_frame);
    // the spans of one learned-PSD prediction each (the whole range without
    // a prediction period), each in launches of equal length
    std::vector<std::pair<int, int>> spans;
    if (learned && learned_period > 0) {
      for (int a = P.a; a <= P.b;) {
        const int b = std::min(P.b, (a / learned_period + 1) * learned_period - 1);
        spans.push_back({a, b});
        a = b + 1;
      }
    } else {
      spans.push_back({P.a, P.b});
    }
    int j = 0;
    long long cur = lspan;
    for (const auto& sp : spans) {
      const int total = sp.second - sp.first + 1;
      const int nl = (total + run_frames - 1) / run_frames;
      for (int jj = 0, m0 = sp.first; jj < nl; jj++, j++) {
        const int nf = total / nl + (jj >= nl - total % nl ? 1 : 0);
        RunLaunch R{};
        R.J = base_job();
        R.J.sigma2 = (float)(P.mode == nf3d::MODE_W3D ? ep.bt : 1) * ep.sigma * ep.sigma / norm;
        R.J.seam = d_seam_run + (size_t)(j & 1) * kRunFrames * seam_frame;
        // learned PSD in d_lpat: predicted from the launch's first frame just
        // before the launch, or, with a prediction period, before the first
        // launch of each span only
        if (learned) use_learned(R.J);
        R.J.src_pitch = src_pitch;
        R.J.dst_pitch = dst_pitch;
        for (int i = 0; i < nf; i++) {
          R.src[i] = static_cast<const uint8_t*>(src[m0 + i + P.to]);
          R.dst[i] = static_cast<uint8_t*>(dst[m0 + i]);
        }
        R.ring = d_ring;
        R.spec_elems = (long long)spec_elems;
        R.m0 = m0;
        R.nframes = nf;
        R.fpw = (P.mode == nf3d::MODE_W3D || P.mode == nf3d::MODE_KALMAN) ? nf : std::min(run_fpw, nf);
        R.ntiles = noy * ntx;
        P.launches.push_back(R);
        const long long idx = learned_period > 0 ? sp.first / learned_period : -1;
        P.predict.push_back(learned_period <= 0 || (jj == 0 && idx != cur));
        if (learned_period > 0) cur = idx;
        m0 += nf;
      }
    }
    if (P.mode == nf3d::MODE_W3D) {
      for (int n = first; n < first + count; n++) {
        if (n >= P.a && n <= P.b) continue;
        RunLaunch R{};
        R.J = base_job();
        R.J.sigma2 = ep.sigma * ep.sigma / norm;  // btcur = 1 (src/fft3d_engine.h's window rule)
        R.J.seam = d_seam_run;  // the edge launches run apart from the others on the stream
        if (learned) use_learned(R.J);
        R.J.src_pitch = src_pitch;
        R.J.dst_pitch = dst_pitch;
        R.src[0] = static_cast<const uint8_t*>(src[n]);
        R.dst[0] = static_cast<uint8_t*>(dst[n]);
        R.m0 = n;
        R.nframes = 1;
        R.fpw = 1;
        R.ntiles = noy * ntx;
        (n < P.a ? P.edges_start : P.edges).push_back(R);
      }
    }
    return P;
  }

  // Frames before the frame-loop launches (or all, without them), the
  // pattern estimate, the Wiener-3D ring slots of the first window.
  void run_prologue(const RangePlan& P, hipStream_t s) {
    check(hipSetDevice(device), "hipSetDevice");
    const int stop = P.run ? P.a : P.first + P.count;
    if (P.edges.empty() && P.edges_start.empty())
      for (int n = P.first; n < stop; n++) run_one(n, P.src, P.src_pitch, P.dst, P.dst_pitch, s);
    if (!P.run) return;
    if (ep.pfactor != 0 && !pattern_set && !learned) {
      const int pf = std::max(0, std::min(ep.pframe, frames - 1));
      nf3d_frame_ref r{pf, P.src[pf], (int)P.src_pitch};
      estimate_pattern(Call{&r, 1, true, s});
    }
    if (P.mode == nf3d::MODE_W3D) {
      if (ep.pfactor != 0 && !learned) ensure_pattern3d(ep.bt, s);
      // the window's older frames of the first step, into slot (frame % bt),
      // by the frame-loop kernel's forward transforms (so every spectrum the
      // loop reads has its arithmetic, whatever the calls' split), unless
      // the slot holds that frame from that kernel already (a previous call)
      int k0 = P.a + P.from;
      while (k0 < P.a + P.to && ring_frame[k0 % ep.bt] == k0 && ring_by_run[k0 % ep.bt]) k0++;
      if (k0 < P.a + P.to) {
        nf3d::RunBatch B{};
        RunLaunch& R = B.job[0];
        R.J = base_job();
        R.J.src_pitch = P.src_pitch;
        R.J.dst_pitch = P.dst_pitch;
        for (int k = k0; k < P.a + P.to; k++) R.src[k - k0] = static_cast<const uint8_t*>(P.src[k]);
        R.ring = d_ring;
        R.spec_elems = (long long)spec_elems;
        R.m0 = k0;
        R.nframes = P.a + P.to - k0;
        R.fpw = R.nframes;
        R.ntiles = noy * ntx;
        R.wg_begin = 0;
        B.njobs = 1;
        B.total_wg = R.ntiles;
        B.max_hc = Hc;
        B.max_frames = R.nframes;
        check(nf3d::launch_run(nf3d::MODE_FWD, ep.bt, nf3d::SHP_NONE, dft_mode, ep.bw, ep.bh, B, s),
              "frame-loop forward transforms");
        for (int k = k0; k < P.a + P.to; k++) {
          ring_frame[k % ep.bt] = k;
          ring_by_run[k % ep.bt] = 1;
          ring_tick[k % ep.bt] = ++tick;
        }
      }
    }
  }

  // Ring bookkeeping after the frame-loop launches, then the frames after them.
  void run_epilogue(const RangePlan& P, hipStream_t s) {
    if (!P.run) return;
    if (P.mode == nf3d::MODE_W3D) {
      // slot (k % bt) now holds frame k for the last bt frames transformed
      std::fill(ring_frame.begin(), ring_frame.end(), -1);
      for (int k = std::max(0, P.b + P.to - ep.bt + 1); k <= P.b + P.to; k++) {
        ring_frame[k % ep.bt] = k;
        ring_by_run[k % ep.bt] = 1;
        ring_tick[k % ep.bt] = ++tick;
      }
    }
    if (P.edges.empty() && P.edges_start.empty())
      for (int n = P.b + 1; n < P.first + P.count; n++) run_one(n, P.src, P.src_pitch, P.dst, P.dst_pitch, s);
  }

  void get_frame_host(int n, const nf3d_frame_ref* frames_, int nframes, void* dst, int dst_stride) {
    Call c{frames_, nframes, false, stream};
    order_before(stream);
    done_stream = nullptr;  // this call waits for its own work below
    std::fill(stage_frame.begin(), stage_frame.end(), -1);
    size_t dpitch = ((size_t)framewidth * pix + 255) / 256 * 256;
    if (!d_dst) d_dst = dev_alloc<uint8_t>(dpitch * frameheight);
    run(n, c, d_dst, (long long)dpitch);
    check(hipMemcpy2DAsync(static_cast<uint8_t*>(dst) + (size_t)tc * dst_stride + (size_t)lc * pix, dst_stride,
                           d_dst + (size_t)tc * dpitch + (size_t)lc * pix, dpitch, (size_t)Wc * pix, Hc,
                           hipMemcpyDeviceToHost, stream),
          "download");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }
};

// ---------------------------------------------------------------- C ABI
extern "C" {

void nf3d_params_defaults(nf3d_params* p, float sigma) {
  // src/fft3d.hpp:77-106
  *p = nf3d_params{};
  p->sigma = sigma;
  p->beta = 1.0f;
  p->bw = 32;
  p->bh = 32;
  p->bt = 3;
  p->ow = -1;
  p->oh = -1;
  p->kratio = 2.0f;
  p->sharpen = 0.0f;
  p->scutoff = 0.3f;
  p->svr = 1.0f;
  p->smin = 4.0f;
  p->smax = 20.0f;
  p->measure = 1;
  p->interlaced = 0;
  p->wintype = 0;
  p->pframe = 0;
  p->px = 0;
  p->py = 0;
  p->pshow = 0;
  p->pcutoff = 0.1f;
  p->pfactor = 0.0f;
  p->sigma2 = sigma;
  p->sigma3 = sigma;
  p->sigma4 = sigma;
  p->degrid = 1.0f;
  p->dehalo = 0.0f;
  p->hr = 2.0f;
  p->ht = 50.0f;
  p->l = p->t = p->r = p->b = 0;
  p->opt = 0;
}

int nf3d_engine_create(const nf3d_params* p, const nf3d_format* fmt, int width, int height, int frames,
                       int plane, nf3d_engine** out, char* err, size_t errlen) {
  if (!p || !fmt || !out) {
    copy_err(err, errlen, "null argument");
    return -1;
  }
  auto* e = new nf3d_engine();
  try {
    e->init(*p, *fmt, width, height, frames, plane);
  } catch (const Failure& f) {
    copy_err(err, errlen, f.msg);
    delete e;
    return -1;
  }
  *out = e;
  return 0;
}

int nf3d_engine_get_frame(nf3d_engine* e, int n, const nf3d_frame_ref* frames, int nframes, void* dst,
                          int dst_stride, char* err, size_t errlen) {
  std::lock_guard<std::mutex> lock(e->mu);
  try {
    e->last_end = -1;
    e->get_frame_host(n, frames, nframes, dst, dst_stride);
  } catch (const Failure& f) {
    copy_err(err, errlen, f.msg);
    return -1;
  }
  return 0;
}

int nf3d_engine_get_frame_device(nf3d_engine* e, int n, const nf3d_frame_ref* frames, int nframes, void* dst,
                                 int dst_stride, void* stream, char* err, size_t errlen) {
  std::lock_guard<std::mutex> lock(e->mu);
  try {
    nf3d_engine::Call c{frames, nframes, true, stream ? static_cast<hipStream_t>(stream) : e->stream};
    e->order_before(c.s);
    e->last_end = -1;
    e->run(n, c, static_cast<uint8_t*>(dst), dst_stride);
    e->order_after(c.s);
  } catch (const Failure& f) {
    copy_err(err, errlen, f.msg);
    return -1;
  }
  return 0;
}

namespace {

// The batch: plans of every item, prologues, the frame-loop launches grouped
// by (block shape, mode, bt, sharpen case, transform form) and launch index,
// epilogues.
// The learned-PSD predictions of one frame-loop launch: the PSD of each
// job's first frame m0 (for Wiener 3D with bt >= 3 its spectrum is in ring
// slot m0 % bt already -- frame m0 is an older frame of its own window -- so
// the prediction starts from there instead of transforming the frame again).
// Jobs that share a model and an output scale (the planes of a clip) are
// predicted together: one feature pass per job into the first one's batch
// buffer, then one set of GEMMs over all their blocks, each job's pattern
// pointer set to its rows -- or, for an engine with a prediction period, its
// rows copied to its own d_lpat, which its later launches of the span reuse.
// `mode` is the launch's kernel mode (the 2D launches of a Wiener-3D range's
// clip-end frames scale by 1, not bt); `l` the launch's index in the jobs'
// plans (SIZE_MAX for the clip-end launches, which all predict).
void learned_predictions(const nf3d_run_item* items, const std::vector<nf3d_engine::RangePlan>& plans,
                         const std::vector<int>& members, size_t c0, nf3d::RunBatch& B, int mode, size_t l,
                         hipStream_t s) {
  std::vector<bool> done(B.njobs, false);
  for (int c = 0; c < B.njobs; c++) {
    const int k = members[c0 + c];
    if (l != SIZE_MAX && !plans[k].predict[l]) done[c] = true;
  }
  for (int c = 0; c < B.njobs; c++) {
    const int k = members[c0 + c];
    nf3d_engine* e = items[k].engine;
    if (!e->learned || done[c]) continue;
    const float mult = mode == nf3d::MODE_W3D ? (float)e->ep.bt : 1.0f;
    const float scale = mult * e->learned_strength;
    std::vector<int> group;
    for (int c2 = c; c2 < B.njobs; c2++) {
      const int k2 = members[c0 + c2];
      nf3d_engine* e2 = items[k2].engine;
      const float mult2 = mode == nf3d::MODE_W3D ? (float)e2->ep.bt : 1.0f;
      if (!done[c2] && e2->learned == e->learned && mult2 * e2->learned_strength == scale) {
        group.push_back(c2);
        done[c2] = true;
      }
    }
    auto spectra = [&](int c2) -> const float2* {
      const int k2 = members[c0 + c2];
      nf3d_engine* e2 = items[k2].engine;
      const RunLaunch& R = B.job[c2];
      if (mode == nf3d::MODE_W3D && e2->ep.bt >= 3) return e2->slot_ptr(R.m0 % e2->ep.bt);
      // a Wiener-3D range's clip-end frame: in the ring when its launch runs
      // (the start ones right after the prologue, the end ones after the
      // frame-loop launches)
      if (l == SIZE_MAX && plans[k2].mode == nf3d::MODE_W3D) return e2->slot_ptr(R.m0 % e2->ep.bt);
      PlaneJob F = e2->base_job();
      F.src = static_cast<const uint8_t*>(plans[k2].src[R.m0]);
      F.src_pitch = plans[k2].src_pitch;
      F.spec_out = e2->d_lspec;
      e2->launch(nf3d::MODE_FWD, 0, F, s);
      return e2->d_lspec;
    };
    // (the clip-end launches predict into the batch buffer even alone, so
    // that d_lpat keeps a span's PSD for the calls that continue it)
    if (group.size() == 1 && l != SIZE_MAX) {
      e->learned_psd_spec(spectra(c), mult, s);
      if (l != SIZE_MAX && e->learned_period > 0) e->lspan = B.job[c].m0 / e->learned_period;
      continue;
    }
    long long rows = 0;
    for (int c2 : group) rows += items[members[c0 + c2]].engine->nb;
    const long long n = (long long)e->ep.bh * e->outwidth;
    const long long work = nf3d_mlp_work_bytes(e->learned, (int)rows);
    if (rows > e->b_rows) {
      if (e->d_bx) e->b_retired.push_back(e->d_bx);
      if (e->d_bp) e->b_retired.push_back(e->d_bp);
      e->d_bx = nullptr;
      e->d_bp = nullptr;
      e->d_bx = dev_alloc<uint16_t>((size_t)(rows * n));
      e->d_bp = dev_alloc<float>((size_t)(rows * n));
      e->b_rows = rows;
    }
    if (work > e->b_work) {
      if (e->d_bwork) e->b_retired.push_back(e->d_bwork);
      e->d_bwork = nullptr;
      e->d_bwork = dev_alloc<uint8_t>((size_t)work);
      e->b_work = work;
    }
    // one feature launch over the group's spectra (at most RUN_MAX_JOBS)
    static_assert(nf3d::RUN_MAX_JOBS <= 8, "the feature launch takes up to 8 spectra arrays");
    const void* specs[nf3d::RUN_MAX_JOBS];
    int seg_rows[nf3d::RUN_MAX_JOBS];
    int nseg = 0;
    long long row = 0;
    for (int c2 : group) {
      nf3d_engine* e2 = items[members[c0 + c2]].engine;
      specs[nseg] = spectra(c2);
      seg_rows[nseg++] = e2->nb;
      if (l == SIZE_MAX || e2->learned_period <= 0) {
        B.job[c2].J.pattern2d = e->d_bp + row * n;
        B.job[c2].J.pattern3d = e->d_bp + row * n;
      }
      row += e2->nb;
    }
    if (nf3d_mlp_features_batch(e->learned, specs, seg_rows, nseg, e->d_bx, s)) fail("learned PSD: feature kernel");
    char msg[256] = {0};
    if (nf3d_mlp_predict(e->learned, e->d_bx, (int)rows, e->d_bp, 1, scale, e->d_bwork, s, msg, sizeof msg))
      fail(std::string("learned PSD: ") + msg);
    row = 0;
    for (int c2 : group) {
      nf3d_engine* e2 = items[members[c0 + c2]].engine;
      if (l != SIZE_MAX && e2->learned_period > 0) {
        check(hipMemcpyAsync(e2->d_lpat, e->d_bp + row * n, sizeof(float) * (size_t)(e2->nb * n),
                             hipMemcpyDeviceToDevice, s),
              "learned PSD: copy");
        e2->lspan = B.job[c2].m0 / e2->learned_period;
      }
      row += e2->nb;
    }
  }
}

// keep_events: under graph capture, the fork and join events are handed
// there, to be destroyed after the capture ends, instead of here.
void run_batch(const nf3d_run_item* items, int n, hipStream_t s, bool timed, bool capture = false,
               std::vector<hipEvent_t>* keep_events = nullptr) {
  if (!capture)
    for (int i = 0; i < n; i++) items[i].engine->order_before(s);
  std::vector<nf3d_engine::RangePlan> plans(n);
  for (int i = 0; i < n; i++) {
    const nf3d_run_item& it = items[i];
    plans[i] = it.engine->plan_range(it.first, it.count, it.src, it.src_stride, it.dst, it.dst_stride);
  }
  // the clip-end frames of Wiener-3D ranges: one-frame launches of the
  // frame-loop kernel in its 2D mode, engines of the same shape, sharpen
  // case and transform form together (those before the frame-loop frames
  // right after the prologues, the others after the frame-loop launches)
  auto run_edges = [&](std::vector<RunLaunch> nf3d_engine::RangePlan::*which) {
    std::vector<bool> edone(n, false);
    for (int i = 0; i < n; i++) {
      if (edone[i] || (plans[i].*which).empty()) continue;
      nf3d_engine* e0 = items[i].engine;
      std::vector<int> group;
      for (int k = i; k < n; k++) {
        nf3d_engine* e = items[k].engine;
        if (!edone[k] && !(plans[k].*which).empty() && e->ep.bw == e0->ep.bw && e->ep.bh == e0->ep.bh &&
            plans[k].shp == plans[i].shp && e->dft_mode == e0->dft_mode) {
          group.push_back(k);
          edone[k] = true;
        }
      }
      size_t nq = 0;
      for (int k : group) nq = std::max(nq, (plans[k].*which).size());
      for (size_t q = 0; q < nq; q++) {
        std::vector<int> members;
        for (int k : group)
          if (q < (plans[k].*which).size()) members.push_back(k);
        for (size_t c0 = 0; c0 < members.size(); c0 += nf3d::RUN_MAX_JOBS) {
          nf3d::RunBatch B{};
          std::vector<std::pair<hipEvent_t, hipEvent_t>> evs;
          for (size_t c = c0; c < members.size() && c < c0 + nf3d::RUN_MAX_JOBS; c++) {
            const int k = members[c];
            nf3d_engine* e = items[k].engine;
            RunLaunch& R = B.job[B.njobs++];
            R = (plans[k].*which)[q];
            R.wg_begin = B.total_wg;
            B.total_wg += R.ntiles;
            B.max_hc = std::max(B.max_hc, e->Hc);
            B.max_frames = 1;
          }
          learned_predictions(items, plans, members, c0, B, nf3d::MODE_W2D, SIZE_MAX, s);
          for (int c = 0; c < B.njobs; c++) evs.push_back(items[members[c0 + c]].engine->timer(timed, 1, s));
          check(nf3d::launch_run(nf3d::MODE_W2D, 0, plans[i].shp, e0->dft_mode, e0->ep.bw, e0->ep.bh, B, s),
                "frame-loop kernel (clip ends)");
          check(nf3d::launch_seam_run(e0->ep.bw, e0->ep.bh, B, s), "seam kernel");
          for (auto& ev : evs)
            if (ev.second) check(hipEventRecord(ev.second, s), "hipEventRecord");
        }
      }
    }
  };
  for (int i = 0; i < n; i++) items[i].engine->run_prologue(plans[i], s);
  run_edges(&nf3d_engine::RangePlan::edges_start);
  // the frame-loop launches, grouped by block shape, mode, bt, sharpen case
  // and transform form
  std::vector<std::vector<int>> groups;
  std::vector<bool> done(n, false);
  for (int i = 0; i < n; i++) {
    if (done[i] || !plans[i].run) continue;
    nf3d_engine* e0 = items[i].engine;
    std::vector<int> group;
    for (int k = i; k < n; k++) {
      nf3d_engine* e = items[k].engine;
      if (!done[k] && plans[k].run && e->ep.bw == e0->ep.bw && e->ep.bh == e0->ep.bh && e->ep.bt == e0->ep.bt &&
          plans[k].mode == plans[i].mode && plans[k].shp == plans[i].shp && e->dft_mode == e0->dft_mode) {
        group.push_back(k);
        done[k] = true;
      }
    }
    groups.push_back(group);
  }
  // groups share no engine: with several, each runs on its first engine's
  // own stream (used by no call but this engine's, which its lock holds
  // off), forked from `s` and joined back to it, so that the launches of one
  // group fill the slots another's launch tails leave (in a captured graph:
  // parallel branches)
  std::vector<hipEvent_t> joins;
  hipEvent_t fork = nullptr;
  struct Events {
    hipEvent_t& fork;
    std::vector<hipEvent_t>& joins;
    std::vector<hipEvent_t>* keep;
    ~Events() {
      if (fork) joins.push_back(fork);
      for (hipEvent_t e : joins) {
        if (keep) keep->push_back(e);
        else (void)hipEventDestroy(e);
      }
    }
  } events{fork, joins, keep_events};
  if (groups.size() > 1) {
    check(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "hipEventCreate");
    check(hipEventRecord(fork, s), "hipEventRecord");
  }
  for (const std::vector<int>& group : groups) {
    const int i = group[0];
    nf3d_engine* e0 = items[i].engine;
    const hipStream_t gs = groups.size() > 1 ? e0->stream : s;
    if (gs != s) check(hipStreamWaitEvent(gs, fork, 0), "hipStreamWaitEvent");
    size_t nl = 0;
    for (int k : group) nl = std::max(nl, plans[k].launches.size());
    for (size_t l = 0; l < nl; l++) {
      std::vector<int> members;
      for (int k : group)
        if (l < plans[k].launches.size()) members.push_back(k);
      for (size_t c0 = 0; c0 < members.size(); c0 += nf3d::RUN_MAX_JOBS) {
        nf3d::RunBatch B{};
        std::vector<std::pair<hipEvent_t, hipEvent_t>> evs;
        for (size_t c = c0; c < members.size() && c < c0 + nf3d::RUN_MAX_JOBS; c++) {
          const int k = members[c];
          nf3d_engine* e = items[k].engine;
          RunLaunch& R = B.job[B.njobs++];
          R = plans[k].launches[l];
          R.wg_begin = B.total_wg;
          B.total_wg += R.ntiles * ((R.nframes + R.fpw - 1) / R.fpw);
          B.max_hc = std::max(B.max_hc, e->Hc);
          B.max_frames = std::max(B.max_frames, R.nframes);
        }
        learned_predictions(items, plans, members, c0, B, plans[i].mode, l, gs);
        for (int c = 0; c < B.njobs; c++)
          evs.push_back(items[members[c0 + c]].engine->timer(timed, B.job[c].nframes, gs));
        // the seam pass follows on the same stream: run beside the next
        // launch it takes issue slots the latency-bound k_run needs more
        check(nf3d::launch_run(plans[i].mode, e0->ep.bt, plans[i].shp, e0->dft_mode, e0->ep.bw, e0->ep.bh, B, gs),
              "frame-loop kernel");
        check(nf3d::launch_seam_run(e0->ep.bw, e0->ep.bh, B, gs), "seam kernel");
        for (auto& ev : evs)
          if (ev.second) check(hipEventRecord(ev.second, gs), "hipEventRecord");
      }
    }
    if (gs != s) {
      hipEvent_t j = nullptr;
      check(hipEventCreateWithFlags(&j, hipEventDisableTiming), "hipEventCreate");
      joins.push_back(j);
      check(hipEventRecord(j, gs), "hipEventRecord");
      check(hipStreamWaitEvent(s, j, 0), "hipStreamWaitEvent");
    }
  }
  run_edges(&nf3d_engine::RangePlan::edges);
  for (int i = 0; i < n; i++) items[i].engine->run_epilogue(plans[i], s);
  for (int i = 0; i < n; i++) items[i].engine->last_end = items[i].first + items[i].count - 1;
  if (!capture)
    for (int i = 0; i < n; i++) items[i].engine->order_after(s);
}

// Locks every engine of a batch (address order), after checking the items.
std::vector<std::unique_lock<std::mutex>> lock_batch(const nf3d_run_item* items, int n) {
  if (n < 1) fail("empty batch");
  std::vector<nf3d_engine*> es;
  for (int i = 0; i < n; i++) {
    if (!items[i].engine) fail("batch item without an engine");
    es.push_back(items[i].engine);
  }
  std::sort(es.begin(), es.end());
  if (std::adjacent_find(es.begin(), es.end()) != es.end()) fail("an engine appears twice in the batch");
  for (nf3d_engine* e : es)
    if (e->device != es[0]->device) fail("batch engines on different devices");
  std::vector<std::unique_lock<std::mutex>> locks;
  for (nf3d_engine* e : es) locks.emplace_back(e->mu);
  check(hipSetDevice(es[0]->device), "hipSetDevice");
  return locks;
}

}  // namespace

int nf3d_engine_run_device(nf3d_engine* e, int first, int count, const void* const* src, int src_stride,
                           void* const* dst, int dst_stride, void* stream, char* err, size_t errlen) {
  std::lock_guard<std::mutex> lock(e->mu);
  try {
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : e->stream;
    check(hipSetDevice(e->device), "hipSetDevice");
    nf3d_run_item it{e, first, count, src, src_stride, dst, dst_stride};
    run_batch(&it, 1, s, true);
  } catch (const Failure& f) {
    copy_err(err, errlen, f.msg);
    return -1;
  }
  return 0;
}


// A captured batch.  The capture runs nothing, but the batch's bookkeeping
// records which frames its launches leave in the spectra ring; those records
// are kept here (per engine) and given to the engines when a replay has
// actually run, while after the capture itself the engines' records say the
// ring holds nothing known.
// --------------------------------------------------------- host-fed ranges
// nf3d_run_host_batch: the batch of nf3d_run_device_batch with its frames in
// host memory, streamed through the device in chunks of kChunk frames.  For
// chunk c the host copies the caller's planes of chunk c+1 into pinned
// staging (worker threads) and queues their upload on the upload stream,
// while the compute stream filters chunk c from device frame slots and the
// download stream returns chunk c-1, whose pinned output the host then copies
// to the caller.  Device slots: input frame k in slot k % kInSlots (a chunk
// plus the Wiener window's frames on either side, and the next chunk being
// uploaded), output frame k in slot k % (2 * kChunk).
namespace {

constexpr int kChunk = 16;
constexpr int kInSlots = 3 * kChunk;

struct HostPlane {
  nf3d_engine* e = nullptr;
  size_t row = 0;      // bytes of one row
  int rows = 0;
  size_t pitch = 0;    // device and pinned slot pitch (256-byte aligned)
  uint8_t* d_in = nullptr;
  uint8_t* d_out = nullptr;
  uint8_t* h_in[2] = {nullptr, nullptr};   // pinned, one chunk each
  uint8_t* h_out[2] = {nullptr, nullptr};
  std::vector<const void*> dsrc;
  std::vector<void*> ddst;
  ~HostPlane() {
    hipFree(d_in);
    hipFree(d_out);
    for (int i = 0; i < 2; i++) {
      hipHostFree(h_in[i]);
      hipHostFree(h_out[i]);
    }
  }
};

// Rows of several planes between strided buffers, split over host threads.
struct CopyJob {
  uint8_t* dst;
  size_t dpitch;
  const uint8_t* src;
  size_t spitch;
  size_t row;
  int rows;
};
// Whether p is page-locked host memory the DMA engines can read directly
// (hipHostMalloc'd; pageable memory is not known to the runtime).
bool is_pinned(const void* p) {
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

// Device slots and pinned staging of an engine's host-fed calls, kept from
// one call to the next and released with the engine (the map itself is never
// destroyed: no HIP call from static destructors after the runtime is gone).
std::mutex g_host_mu;
auto& g_host_planes = *new std::unordered_map<const nf3d_engine*, std::unique_ptr<HostPlane>>;

void run_copies(const std::vector<CopyJob>& jobs) {
  if (jobs.empty()) return;
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const size_t nt = std::min<size_t>(hw, jobs.size());
  auto work = [&](size_t t) {
    for (size_t j = t; j < jobs.size(); j += nt) {
      const CopyJob& c = jobs[j];
      if (c.dpitch == c.spitch && c.dpitch == c.row) {
        memcpy(c.dst, c.src, c.row * c.rows);
      } else {
        for (int y = 0; y < c.rows; y++) memcpy(c.dst + y * c.dpitch, c.src + y * c.spitch, c.row);
      }
    }
  };
  std::vector<std::thread> th;
  for (size_t t = 1; t < nt; t++) th.emplace_back(work, t);
  work(0);
  for (auto& t : th) t.join();
}

}  // namespace

struct nf3d_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<nf3d_engine*> engines;
  std::vector<std::vector<int>> ring_after;  // ring_frame of each engine after one replay
  std::vector<std::vector<char>> by_run_after;
  // captured against the rings as they stood (nf3d_run_device_batch_graph_ring):
  // what each ring held then; a replay needs the same frames, shifted by a
  // whole number of frames per engine
  bool ring_bound = false;
  std::vector<std::vector<int>> ring_before;
  std::vector<std::vector<char>> by_run_before;
  // each engine's last frame (last_end) and learned-PSD span (lspan) at
  // capture and after one replay
  std::vector<int> end_before, end_after;
  std::vector<long long> lspan_before, lspan_after;
};

int nf3d_run_host_batch(const nf3d_run_item* items, int n, char* err, size_t errlen) {
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  try {
    auto locks = lock_batch(items, n);
    int first = items[0].first, count = items[0].count;
    for (int i = 0; i < n; i++) {
      const nf3d_run_item& it = items[i];
      if (it.first != first || it.count != count) fail("host batch items must share one frame range");
      if (!it.src || !it.dst) fail("null frame list");
      for (int k = std::max(0, first - 2); k < std::min(it.engine->frames, first + count + 2); k++)
        if (!it.src[k] || (k >= first && k < first + count && !it.dst[k]))
          fail("frame " + std::to_string(k) + " was not provided");
    }
    if (count < 1) return 0;
    const int F = items[0].engine->frames;
    for (int i = 0; i < n; i++)
      if (items[i].engine->frames != F) fail("host batch engines must have the same clip length");
    for (int k = 0; k < 3; k++) {
      hipStream_t st;
      check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
      streams.push_back(st);
    }
    hipStream_t s_up = streams[0], s_run = streams[1], s_down = streams[2];
    // per chunk: uploaded, filtered, downloaded
    const int nchunks = (count + kChunk - 1) / kChunk;
    auto new_event = [&]() {
      hipEvent_t ev;
      check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
      events.push_back(ev);
      return ev;
    };
    std::vector<hipEvent_t> ev_up(nchunks + 1), ev_run(nchunks), ev_down(nchunks);
    for (auto& e : ev_up) e = new_event();
    for (auto& e : ev_run) e = new_event();
    for (auto& e : ev_down) e = new_event();
    std::vector<HostPlane*> pl;
    std::vector<char> pin_src(n), pin_dst(n);  // the caller's planes are pinned: no staging copy
    for (int i = 0; i < n; i++) {
      nf3d_engine* e = items[i].engine;
      std::unique_ptr<HostPlane>* cached;
      {
        std::lock_guard<std::mutex> lock(g_host_mu);
        cached = &g_host_planes[e];
      }
      if (!*cached) {
        auto P = std::make_unique<HostPlane>();
        P->e = e;
        P->row = (size_t)e->framewidth * e->pix;
        P->rows = e->frameheight;
        // rows that are a multiple of 64 bytes keep their width as the slot
        // pitch, so a frame (and a run of frames) is one contiguous copy
        P->pitch = P->row % 64 == 0 ? P->row : (P->row + 255) / 256 * 256;
        const size_t slot = P->pitch * P->rows;
        P->d_in = dev_alloc<uint8_t>(slot * kInSlots);
        P->d_out = dev_alloc<uint8_t>(slot * 2 * kChunk);
        for (int b = 0; b < 2; b++) {
          check(hipHostMalloc((void**)&P->h_in[b], slot * (kChunk + 4), hipHostMallocDefault), "hipHostMalloc");
          check(hipHostMalloc((void**)&P->h_out[b], slot * kChunk, hipHostMallocDefault), "hipHostMalloc");
        }
        P->dsrc.assign(F, nullptr);
        P->ddst.assign(F, nullptr);
        for (int k = 0; k < F; k++) {
          P->dsrc[k] = P->d_in + (size_t)(k % kInSlots) * slot;
          P->ddst[k] = P->d_out + (size_t)(k % (2 * kChunk)) * slot;
        }
        *cached = std::move(P);
      }
      pl.push_back(cached->get());
      bool ps = true, pd = true;
      for (int k = std::max(0, first - 2); k < std::min(F, first + count + 2) && ps; k++) ps = is_pinned(items[i].src[k]);
      for (int k = first; k < first + count && pd; k++) pd = is_pinned(items[i].dst[k]);
      pin_src[i] = ps;
      pin_dst[i] = pd;
    }
    for (int i = 0; i < n; i++) items[i].engine->order_before(s_up);
    // frames a chunk's filter reads: its own and the window's on either side
    auto chunk_range = [&](int c, int& lo, int& hi) {
      const int a = first + c * kChunk, b = std::min(first + count, a + kChunk) - 1;
      lo = std::max(0, a - 2);
      hi = std::min(F - 1, b + 2);
    };
    // frames chunk c uploads: those of its range not uploaded by chunk c-1
    auto upload_range = [&](int c, int& lo, int& hi) {
      chunk_range(c, lo, hi);
      if (c > 0) {
        int plo, phi;
        chunk_range(c - 1, plo, phi);
        lo = std::max(lo, phi + 1);
      }
    };
    auto upload = [&](int c) {
      int lo, hi;
      upload_range(c, lo, hi);
      const int buf = c & 1;
      if (c >= 2) check(hipEventSynchronize(ev_up[c - 2]), "hipEventSynchronize");  // staging free again
      std::vector<CopyJob> jobs;
      for (int i = 0; i < n; i++) {
        HostPlane& P = *pl[i];
        if (pin_src[i]) continue;
        for (int k = lo; k <= hi; k++)
          jobs.push_back(CopyJob{P.h_in[buf] + (size_t)(k - lo) * P.pitch * P.rows, P.pitch,
                                 static_cast<const uint8_t*>(items[i].src[k]), (size_t)items[i].src_stride, P.row,
                                 P.rows});
      }
      run_copies(jobs);
      // slots of frames the chunk two back still read are free once it ran
      if (c >= 2) check(hipStreamWaitEvent(s_up, ev_run[c - 2], 0), "hipStreamWaitEvent");
      for (int i = 0; i < n; i++) {
        HostPlane& P = *pl[i];
        const size_t frame = P.pitch * P.rows;
        const bool dense = pin_src[i] && (size_t)items[i].src_stride == P.pitch;
        for (int k = lo; k <= hi;) {
          // a run of frames adjacent both in the source and in the device
          // slots goes up as one copy
          int k1 = k + 1;
          if (!pin_src[i] || dense) {
            while (k1 <= hi && (k1 % kInSlots) != 0 &&
                   (!pin_src[i] || static_cast<const uint8_t*>(items[i].src[k1]) ==
                                       static_cast<const uint8_t*>(items[i].src[k1 - 1]) + frame))
              k1++;
          }
          if (pin_src[i] && !dense)
            check(hipMemcpy2DAsync(const_cast<void*>(P.dsrc[k]), P.pitch, items[i].src[k], (size_t)items[i].src_stride,
                                   P.row, P.rows, hipMemcpyHostToDevice, s_up),
                  "upload");
          else
            check(hipMemcpyAsync(const_cast<void*>(P.dsrc[k]),
                                 pin_src[i] ? items[i].src[k] : P.h_in[buf] + (size_t)(k - lo) * frame,
                                 frame * (size_t)(k1 - k), hipMemcpyHostToDevice, s_up),
                  "upload");
          k = k1;
        }
      }
      check(hipEventRecord(ev_up[c], s_up), "hipEventRecord");
    };
    auto download_finish = [&](int c) {
      check(hipEventSynchronize(ev_down[c]), "hipEventSynchronize");
      const int a = first + c * kChunk, b = std::min(first + count, a + kChunk) - 1;
      std::vector<CopyJob> jobs;
      for (int i = 0; i < n; i++) {
        HostPlane& P = *pl[i];
        nf3d_engine* e = P.e;
        if (pin_dst[i]) continue;  // downloaded straight into the caller's planes
        // the processed region only: crop borders stay the caller's
        for (int k = a; k <= b; k++)
          jobs.push_back(CopyJob{static_cast<uint8_t*>(items[i].dst[k]) + (size_t)e->tc * items[i].dst_stride +
                                     (size_t)e->lc * e->pix,
                                 (size_t)items[i].dst_stride,
                                 P.h_out[c & 1] + (size_t)(k - a) * P.pitch * P.rows + (size_t)e->tc * P.pitch +
                                     (size_t)e->lc * e->pix,
                                 P.pitch, (size_t)e->Wc * e->pix, e->Hc});
      }
      run_copies(jobs);
    };
    upload(0);
    for (int c = 0; c < nchunks; c++) {
      const int a = first + c * kChunk, cnt = std::min(first + count, a + kChunk) - a;
      // filter chunk c once its frames are up
      check(hipStreamWaitEvent(s_run, ev_up[c], 0), "hipStreamWaitEvent");
      if (c >= 2) check(hipStreamWaitEvent(s_run, ev_down[c - 2], 0), "hipStreamWaitEvent");  // output slots
      std::vector<nf3d_run_item> sub(n);
      for (int i = 0; i < n; i++)
        sub[i] = nf3d_run_item{items[i].engine, a, cnt, pl[i]->dsrc.data(), (int)pl[i]->pitch, pl[i]->ddst.data(),
                               (int)pl[i]->pitch};
      run_batch(sub.data(), n, s_run, true);
      check(hipEventRecord(ev_run[c], s_run), "hipEventRecord");
      // return it
      check(hipStreamWaitEvent(s_down, ev_run[c], 0), "hipStreamWaitEvent");
      for (int i = 0; i < n; i++) {
        HostPlane& P = *pl[i];
        nf3d_engine* e = P.e;
        const size_t frame = P.pitch * P.rows;
        // page-locked destination without crop and with the slots' pitch:
        // whole frames, a run of adjacent ones per copy
        const bool dense = pin_dst[i] && (size_t)items[i].dst_stride == P.pitch && e->Wc == e->framewidth &&
                           e->Hc == e->frameheight;
        if (!pin_dst[i] || dense) {
          // the chunk's output slots are adjacent (slot k % (2 kChunk), chunks start at multiples of kChunk
          // from `first`, so a run stops where the slot index wraps)
          for (int k = a; k < a + cnt;) {
            int k1 = k + 1;
            while (k1 < a + cnt && (k1 % (2 * kChunk)) != 0 &&
                   (!pin_dst[i] ||
                    static_cast<uint8_t*>(items[i].dst[k1]) == static_cast<uint8_t*>(items[i].dst[k1 - 1]) + frame))
              k1++;
            check(hipMemcpyAsync(pin_dst[i] ? items[i].dst[k] : P.h_out[c & 1] + (size_t)(k - a) * frame, P.ddst[k],
                                 frame * (size_t)(k1 - k), hipMemcpyDeviceToHost, s_down),
                  "download");
            k = k1;
          }
          continue;
        }
        for (int k = a; k < a + cnt; k++)
          check(hipMemcpy2DAsync(static_cast<uint8_t*>(items[i].dst[k]) + (size_t)e->tc * items[i].dst_stride +
                                     (size_t)e->lc * e->pix,
                                 (size_t)items[i].dst_stride,
                                 static_cast<const uint8_t*>(P.ddst[k]) + (size_t)e->tc * P.pitch +
                                     (size_t)e->lc * e->pix,
                                 P.pitch, (size_t)e->Wc * e->pix, e->Hc, hipMemcpyDeviceToHost, s_down),
                "download");
      }
      check(hipEventRecord(ev_down[c], s_down), "hipEventRecord");
      // meanwhile: the next chunk's frames up, the previous chunk's out
      if (c + 1 < nchunks) upload(c + 1);
      if (c >= 1) download_finish(c - 1);
    }
    download_finish(nchunks - 1);
    check(hipStreamSynchronize(s_up), "hipStreamSynchronize");
    for (int i = 0; i < n; i++) items[i].engine->done_stream = nullptr;  // this call waited for its work
  } catch (const Failure& f) {
    for (auto st : streams) (void)hipStreamSynchronize(st);
    for (auto ev : events) (void)hipEventDestroy(ev);
    for (auto st : streams) (void)hipStreamDestroy(st);
    copy_err(err, errlen, f.msg);
    return -1;
  }
  for (auto ev : events) (void)hipEventDestroy(ev);
  for (auto st : streams) (void)hipStreamDestroy(st);
  return 0;
}

int nf3d_run_device_batch(const nf3d_run_item* items, int n, void* stream, char* err, size_t errlen) {
  try {
    auto locks = lock_batch(items, n);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : items[0].engine->stream;
    run_batch(items, n, s, true);
  } catch (const Failure& f) {
    copy_err(err, errlen, f.msg);
    return -1;
  }
  return 0;
}

namespace {
int capture_batch(const nf3d_run_item* items, int n, void* stream, nf3d_graph** graph, bool keep_ring, char* err,
                  size_t errlen) {
  *graph = nullptr;
  hipStream_t s = static_cast<hipStream_t>(stream);
  bool capturing = false;
  // fork and join events of the captured batch, destroyed once the capture
  // has ended (on return)
  struct CapEvents {
    std::vector<hipEvent_t> v;
    ~CapEvents() {
      for (hipEvent_t e : v) (void)hipEventDestroy(e);
    }
  } cap;
  std::vector<hipEvent_t>& cap_events = cap.v;
  try {
    if (!s) fail("graph capture needs a stream");
    auto locks = lock_batch(items, n);
    for (int i = 0; i < n; i++) {
      nf3d_engine* e = items[i].engine;
      if (e->ep.pfactor != 0 && e->ep.pshow) {
        if (!e->d_ones)
          fail("the pattern display allocates its buffers on its first call, which cannot be captured: run the "
               "batch once first");
      } else if (e->ep.pfactor != 0 && !e->learned &&
                 (!e->pattern_set || (e->ep.bt >= 2 && !e->pattern3d_init))) {
        fail("the noise pattern is estimated on the first call, which cannot be captured: run the batch once first");
      }
      // buffers the launches use, allocated outside the capture
      e->plan_range(items[i].first, items[i].count, items[i].src, items[i].src_stride, items[i].dst,
                    items[i].dst_stride);
    }
    // no profiling events inside the graph
    std::vector<bool> prof(n);
    for (int i = 0; i < n; i++) {
      prof[i] = items[i].engine->profiling;
      items[i].engine->profiling = false;
    }
    struct Restore {
      const nf3d_run_item* items;
      std::vector<bool>& prof;
      ~Restore() {
        for (size_t i = 0; i < prof.size(); i++) items[i].engine->profiling = prof[i];
      }
    } restore{items, prof};
    // a graph must not depend on what the spectra ring holds when it is
    // replayed: forget its contents, so the captured work transforms every
    // frame it reads -- unless it is bound to the rings as they stand (then
    // the captured work reads the frames they hold, and a replay checks them)
    std::vector<std::vector<int>> before(n);
    std::vector<std::vector<char>> by_before(n);
    std::vector<int> end_before(n);
    std::vector<long long> lspan_before(n);
    for (int i = 0; i < n; i++) {
      nf3d_engine* e = items[i].engine;
      before[i] = e->ring_frame;
      by_before[i] = e->ring_by_run;
      end_before[i] = e->last_end;
      lspan_before[i] = e->lspan;
      if (!keep_ring) {
        std::fill(e->ring_frame.begin(), e->ring_frame.end(), -1);
        e->lspan = -1;  // and predicts the learned PSD itself
      }
    }
    check(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    capturing = true;
    try {
      run_batch(items, n, s, false, true, &cap_events);
    } catch (const Failure&) {
      // a capture cut short may have recorded frames in the rings or a kept
      // learned PSD that nothing computed: forget them (the locks still held)
      for (int i = 0; i < n; i++) {
        nf3d_engine* e = items[i].engine;
        std::fill(e->ring_frame.begin(), e->ring_frame.end(), -1);
        e->lspan = -1;
        e->last_end = -1;
      }
      throw;
    }
    auto* g = new nf3d_graph;
    capturing = false;
    g->ring_bound = keep_ring;
    // what a replay leaves in each ring; nothing has run yet
    for (int i = 0; i < n; i++) {
      nf3d_engine* e = items[i].engine;
      g->engines.push_back(e);
      g->ring_after.push_back(e->ring_frame);
      g->by_run_after.push_back(e->ring_by_run);
      g->end_before.push_back(end_before[i]);
      g->end_after.push_back(e->last_end);
      g->lspan_before.push_back(lspan_before[i]);
      g->lspan_after.push_back(e->lspan);
      e->last_end = end_before[i];
      e->lspan = lspan_before[i];
      if (keep_ring) {
        e->ring_frame = before[i];
        e->ring_by_run = by_before[i];
      } else {
        std::fill(e->ring_frame.begin(), e->ring_frame.end(), -1);
      }
    }
    if (keep_ring) {
      g->ring_before = std::move(before);
      g->by_run_before = std::move(by_before);
    }
    hipError_t e1 = hipStreamEndCapture(s, &g->graph);
    if (e1 != hipSuccess) {
      delete g;
      check(e1, "hipStreamEndCapture");
    }
    hipError_t e2 = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e2 != hipSuccess) {
      (void)hipGraphDestroy(g->graph);
      delete g;
      check(e2, "hipGraphInstantiate");
    }
    *graph = g;
  } catch (const Failure& f) {
    if (capturing) {
      hipGraph_t dead = nullptr;
      if (hipStreamEndCapture(s, &dead) == hipSuccess && dead) (void)hipGraphDestroy(dead);
    }
    copy_err(err, errlen, f.msg);
    return -1;
  }
  return 0;
}

// The frames by which each engine has moved on since a ring-bound graph's
// capture (its last frame then and now), if what the captured work reads
// is there: every ring slot the capture saw holds its frame shifted by that
// many frames, with the same kernel-written flag (a multiple of bt: the
// frame loop's slot of frame f is f mod bt), and a learned PSD the capture
// reused is the one of the span shifted alike.  False otherwise.
bool replay_shift(const nf3d_graph* g, std::vector<int>& shift) {
  shift.assign(g->engines.size(), 0);
  for (size_t i = 0; i < g->engines.size(); i++) {
    const nf3d_engine* e = g->engines[i];
    const std::vector<int>& was = g->ring_before[i];
    bool relies = g->lspan_before[i] >= 0;
    for (int f : was) relies = relies || f >= 0;
    const bool known = e->last_end >= 0 && g->end_before[i] >= 0;
    const int d = known ? e->last_end - g->end_before[i] : 0;
    shift[i] = d;
    if (!relies) continue;
    if (!known || e->ring_frame.size() != was.size()) return false;
    if (e->ep.bt >= 2 && d % e->ep.bt != 0) return false;
    for (size_t k = 0; k < was.size(); k++)
      if (was[k] >= 0 && (e->ring_frame[k] != was[k] + d || e->ring_by_run[k] != g->by_run_before[i][k]))
        return false;
    if (g->lspan_before[i] >= 0) {
      const int per = e->learned_period;
      if (per <= 0 || d % per != 0 || e->lspan != g->lspan_before[i] + d / per) return false;
    }
  }
  return true;
}
}  // namespace

int nf3d_run_device_batch_graph(const nf3d_run_item* items, int n, void* stream, nf3d_graph** graph, char* err,
                                size_t errlen) {
  return capture_batch(items, n, stream, graph, false, err, errlen);
}

int nf3d_run_device_batch_graph_ring(const nf3d_run_item* items, int n, void* stream, nf3d_graph** graph,
                                     char* err, size_t errlen) {
  return capture_batch(items, n, stream, graph, true, err, errlen);
}

int nf3d_graph_launch(nf3d_graph* g, void* stream) {
  if (!g || !g->exec) return -1;
  std::vector<nf3d_engine*> es = g->engines;
  std::sort(es.begin(), es.end());
  std::vector<std::unique_lock<std::mutex>> locks;
  for (nf3d_engine* e : es) locks.emplace_back(e->mu);
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::vector<int> shift(g->engines.size(), 0);
  if (g->ring_bound && !replay_shift(g, shift)) return -1;
  try {
    for (nf3d_engine* e : g->engines) e->order_before(s);
    check(hipGraphLaunch(g->exec, s), "hipGraphLaunch");
    for (nf3d_engine* e : g->engines) e->order_after(s);
  } catch (const Failure&) {
    return -1;
  }
  // the replay leaves the frames recorded at capture in the rings (moved on
  // by the rings' shift, for a ring-bound graph)
  for (size_t i = 0; i < g->engines.size(); i++) {
    nf3d_engine* e = g->engines[i];
    e->ring_frame = g->ring_after[i];
    for (int& f : e->ring_frame)
      if (f >= 0) f += shift[i];
    e->ring_by_run = g->by_run_after[i];
    e->last_end = g->end_after[i] + shift[i];
    // the kept span moves with the frames, when they moved by whole spans
    e->lspan = g->lspan_after[i];
    if (e->lspan >= 0) {
      const int per = e->learned_period;
      e->lspan = per > 0 && shift[i] % per == 0 ? e->lspan + shift[i] / per : -1;
    }
  }
  return 0;
}

void nf3d_graph_destroy(nf3d_graph* g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
}

int nf3d_dft2d_mfma(const void* in, void* out, int nblocks, int n, int inverse, float scale, void* stream) {
  return nf3d::launch_dft_mfma(in, out, nblocks, n, inverse, scale, static_cast<hipStream_t>(stream)) == hipSuccess
             ? 0
             : -1;
}

int nf3d_engine_spectra(nf3d_engine* e, const void* src, int src_stride, void* spec_out, void* stream, char* err,
                        size_t errlen) {
  std::lock_guard<std::mutex> lock(e->mu);
  try {
    if (!src || !spec_out) fail("null plane or spectrum");
    check(hipSetDevice(e->device), "hipSetDevice");
    PlaneJob F = e->base_job();
    F.src = static_cast<const uint8_t*>(src);
    F.src_pitch = src_stride;
    F.spec_out = static_cast<float2*>(spec_out);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : e->stream;
    e->order_before(s);
    e->launch(nf3d::MODE_FWD, 0, F, s);
    e->order_after(s);
  } catch (const Failure& f) {
    copy_err(err, errlen, f.msg);
    return -1;
  }
  return 0;
}

int nf3d_engine_region_psd(nf3d_engine* e, const void* src, int src_stride, int rx, int ry, float* out,
                           void* stream, char* err, size_t errlen) {
  std::lock_guard<std::mutex> lock(e->mu);
  try {
    if (!src || !out) fail("null plane or output");
    if (rx < 1 || ry < 1) fail("pattern regions: both sizes >= 1 blocks");
    check(hipSetDevice(e->device), "hipSetDevice");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : e->stream;
    e->order_before(s);
    e->region_psd(static_cast<const uint8_t*>(src), src_stride, rx, ry, out, s);
    e->order_after(s);
  } catch (const Failure& f) {
    copy_err(err, errlen, f.msg);
    return -1;
  }
  return 0;
}

int nf3d_engine_set_learned(nf3d_engine* e, nf3d_mlp* m, float strength, char* err, size_t errlen) {
  std::lock_guard<std::mutex> lock(e->mu);
  try {
    check(hipSetDevice(e->device), "hipSetDevice");
    if (!m) {
      e->learned = nullptr;
      return 0;
    }
    int d[3];
    nf3d_mlp_dims(m, d);
    const int want = e->ep.bh * e->outwidth;
    if (d[0] != want || d[2] != want)
      fail("learned PSD: the model must have " + std::to_string(want) + " inputs and outputs");
    if (!(strength > 0.0f)) fail("learned PSD: strength must be positive");
    if (!e->d_lspec) {
      e->d_lspec = dev_alloc<float2>(e->spec_elems);
      e->d_lpat = dev_alloc<float>(e->spec_elems);
      e->d_lx = dev_alloc<uint16_t>(e->spec_elems);
    }
    hipFree(e->d_lwork);
    e->d_lwork = nullptr;
    e->d_lwork = dev_alloc<uint8_t>((size_t)nf3d_mlp_work_bytes(m, e->nb));
    e->learned = m;
    e->learned_strength = strength;
    e->lspan = -1;
    // the learned PSD is predicted once per frame-loop launch, so frames per
    // launch is part of its result: a tuning-table entry (chosen for speed)
    // does not apply to it; only nf3d_engine_set_tuning changes it
    if (e->run_frames_from_table) {
      e->run_frames = nf3d_engine::kRunFrames;
      e->run_frames_from_table = false;
    }
  } catch (const Failure& f) {
    copy_err(err, errlen, f.msg);
    return -1;
  }
  return 0;
}

int nf3d_engine_set_learned_period(nf3d_engine* e, int frames, char* err, size_t errlen) {
  std::lock_guard<std::mutex> lock(e->mu);
  try {
    if (frames < 0) fail("learned PSD: the prediction period must be 0 (one per launch) or a frame count");
    e->learned_period = frames;
    e->lspan = -1;
  } catch (const Failure& f) {
    copy_err(err, errlen, f.msg);
    return -1;
  }
  return 0;
}

int nf3d_engine_set_pattern_regions(nf3d_engine* e, int rx, int ry, char* err, size_t errlen) {
  std::lock_guard<std::mutex> lock(e->mu);
  try {
    if (rx < 0 || ry < 0 || (rx == 0) != (ry == 0)) fail("pattern regions: both sizes >= 1 blocks, or both 0");
    const bool sigmas = e->ep.sigma2 != e->ep.sigma || e->ep.sigma3 != e->ep.sigma || e->ep.sigma4 != e->ep.sigma;
    if (rx > 0 && (e->ep.pfactor == 0 || e->ep.pshow))
      fail("pattern regions need the estimated noise pattern (pfactor != 0, no pattern display)");
    if (rx > 0 && sigmas && e->pattern_set && !e->d_psums)
      fail("pattern regions need the estimated noise pattern, not one from sigma2..sigma4");
    check(hipSetDevice(e->device), "hipSetDevice");
    if (rx > 0 && !e->d_rpat2d) {
      e->d_rpat2d = dev_alloc<float>(e->spec_elems);
      e->d_rpat3d = dev_alloc<float>(e->spec_elems);
    }
    e->preg_x = rx;
    e->preg_y = ry;
    // estimate again (on the next call) with the new regions
    if (e->d_psums) {
      e->pattern_set = false;
      e->pattern3d_init = false;
    }
  } catch (const Failure& f) {
    copy_err(err, errlen, f.msg);
    return -1;
  }
  return 0;
}

int nf3d_engine_set_dft(nf3d_engine* e, int mode, char* err, size_t errlen) {
  std::lock_guard<std::mutex> lock(e->mu);
  try {
    if (mode != 0 && mode != 1) fail("dft mode must be 0 (FFT) or 1 (matrix-core DFT)");
    if (mode == 1 && !nf3d::mfma_dft_supported(e->ep.bw, e->ep.bh))
      fail("the matrix-core DFT supports 32x32, 48x48 and 64x64 blocks");
    // operands are scaled to FP16 from the format's range, which float
    // samples do not bound
    if (mode == 1 && e->fmt.is_float) fail("the matrix-core DFT supports integer formats only");
    check(hipSetDevice(e->device), "hipSetDevice");
    if (mode == 1 && !e->dft_tab) {
      hipError_t he;
      e->dft_tab = nf3d::mfma_dft_table(e->ep.bw, &he);
      check(he, "matrix-core DFT tables");
    }
    // power-of-two operand scales keeping every FP16 operand below 2^14:
    // windowed samples up to the format's range (float: 4), spectra up to
    // bw*bh times that with headroom 4 for the gains
    const double maxabs = e->fmt.is_float ? 4.0 : (double)((1 << e->fmt.bits_per_sample) - 1);
    const double area = (double)e->ep.bw * e->ep.bh;
    e->dft_sx = std::ldexp(1.0f, (int)std::floor(std::log2(16384.0 / maxabs)));
    e->dft_sz = std::ldexp(1.0f, (int)std::floor(std::log2(16384.0 / (area * maxabs * 4.0))));
    e->dft_mode = mode;
  } catch (const Failure& f) {
    copy_err(err, errlen, f.msg);
    return -1;
  }
  return 0;
}

int nf3d_tune_load(const char* path, char* err, size_t errlen) {
  std::vector<TuneEntry> t;
  std::string msg;
  if (!path) {
    copy_err(err, errlen, "null path");
    return -1;
  }
  int n = parse_tune_file(path, t, msg);
  if (n < 0) {
    copy_err(err, errlen, msg);
    return -1;
  }
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune = std::move(t);
  g_tune_env_done = true;
  return n;
}

int nf3d_tune_lookup(int bw, int bh, int bits, int bt, int* out2) {
  TuneEntry t;
  if (!tune_lookup(bw, bh, bits, bt, &t)) return 0;
  out2[0] = t.frames;
  out2[1] = t.fpw;
  return 1;
}

int nf3d_engine_set_tuning(nf3d_engine* e, int frames_per_launch, int frames_per_workgroup, char* err,
                           size_t errlen) {
  if (frames_per_launch < 1 || frames_per_launch > nf3d::RUN_MAX_FRAMES || frames_per_workgroup < 1 ||
      frames_per_workgroup > nf3d::RUN_MAX_FRAMES) {
    copy_err(err, errlen, "frames per launch and per workgroup must be 1.." + std::to_string(nf3d::RUN_MAX_FRAMES));
    return -1;
  }
  std::lock_guard<std::mutex> lock(e->mu);
  e->run_frames = frames_per_launch;
  e->run_fpw = frames_per_workgroup;
  e->run_frames_from_table = false;
  return 0;
}

int nf3d_engine_get_tuning(nf3d_engine* e, int* out2) {
  std::lock_guard<std::mutex> lock(e->mu);
  out2[0] = e->run_frames;
  out2[1] = e->run_fpw;
  return 0;
}

int nf3d_engine_cache_refresh(nf3d_engine* e, int n) {
  std::lock_guard<std::mutex> lock(e->mu);
  if (e->ring_slots == 0) return 0;
  return e->ring_lookup(n) >= 0 ? 1 : 0;
}

int nf3d_engine_geometry(const nf3d_engine* e, int* o) {
  o[0] = e->nox;
  o[1] = e->noy;
  o[2] = e->outwidth;
  o[3] = e->mirw;
  o[4] = e->mirh;
  o[5] = e->ep.ow;
  o[6] = e->ep.oh;
  return 0;
}

int nf3d_engine_profile(nf3d_engine* e, int enable) {
  std::lock_guard<std::mutex> lock(e->mu);
  e->profiling = enable != 0;
  e->ev_used = 0;
  return 0;
}

int nf3d_engine_profile_read(nf3d_engine* e, double* total_ms, int* frames) {
  std::lock_guard<std::mutex> lock(e->mu);
  try {
    *total_ms = e->profile_read(frames);
  } catch (const Failure&) {
    return -1;
  }
  return 0;
}

int nf3d_engine_sync(nf3d_engine* e) {
  std::lock_guard<std::mutex> lock(e->mu);
  if (hipStreamSynchronize(e->stream) != hipSuccess) return -1;
  if (e->done_ev && hipEventSynchronize(e->done_ev) != hipSuccess) return -1;
  return 0;
}

void nf3d_engine_destroy(nf3d_engine* e) {
  {
    std::lock_guard<std::mutex> lock(g_host_mu);
    g_host_planes.erase(e);
  }
  delete e;
}

}  // extern "C"
