// Test infrastructure only (the checker, never the product).
//
// A minimal in-memory host that drives the reference's own, unmodified
// FFT3D front end (src/fft3d.hpp:12-364) and FFT3DEngine
// (src/fft3d_engine.h:22-741), compiled from the reference tree by
// oracle/build_ref.sh into oracle/_ref/libnf3d_ref.so.
//
// It stands in for the VapourSynth core at the points the DualSynth bridge
// touches (include/dualsynth/ds_frame.hpp:35-146, vs_wrapper.hpp:19-164):
// frames are malloc'd planes behind a small VSAPI function table, and the
// filter parameters come from a "name=value;..." string through an
// InDelegator with VapourSynth semantics (absent key keeps the default,
// vs_wrapper.hpp:23-27).  Nothing here is shipped or measured as the product;
// it produces golden vectors and the CPU-baseline timing of the reference.

#include "fft3d_common.h"
#include <ds_filter.hpp>
#include "fft3d.hpp"

#include <atomic>
#include <mutex>
#include <thread>
#include <cstdlib>
#include <map>
#include <memory>
#include <new>
#include <sstream>
#include <string>
#include <vector>

const AVS_Linkage* AVS_linkage = nullptr;  // no AviSynth host: PVideoFrame calls are no-ops

struct VSFrameRef {
  const VSFormat* fmt{nullptr};
  int width[4]{}, height[4]{}, stride[4]{};
  const uint8_t* rptr[4]{};
  uint8_t* wptr[4]{};
  uint8_t* owned{nullptr};
  std::atomic<int> refs{1};
  bool external{false};  // caller memory: never freed here
};

namespace {

void plane_dims(const VSFormat* f, int w, int h, int p, int& pw, int& ph) {
  bool sub = f->colorFamily == cmYUV && p > 0 && p < 3;
  pw = sub ? (w >> f->subSamplingW) : w;
  ph = sub ? (h >> f->subSamplingH) : h;
}

const VSFrameRef* VS_CC fake_clone(const VSFrameRef* f) noexcept {
  const_cast<VSFrameRef*>(f)->refs.fetch_add(1);
  return f;
}

void VS_CC fake_free(const VSFrameRef* f) noexcept {
  if (!f) return;
  auto* m = const_cast<VSFrameRef*>(f);
  if (m->refs.fetch_sub(1) == 1 && !m->external) {
    free(m->owned);
    delete m;
  }
}

VSFrameRef* alloc_frame(const VSFormat* fmt, int w, int h) {
  auto* fr = new VSFrameRef();
  fr->fmt = fmt;
  size_t total = 0;
  size_t offs[4];
  for (int p = 0; p < fmt->numPlanes; p++) {
    plane_dims(fmt, w, h, p, fr->width[p], fr->height[p]);
    fr->stride[p] = ((fr->width[p] * fmt->bytesPerSample + 63) / 64) * 64;
    offs[p] = total;
    total += size_t(fr->stride[p]) * fr->height[p];
  }
  fr->owned = static_cast<uint8_t*>(aligned_alloc(64, ((total + 63) / 64) * 64));
  memset(fr->owned, 0, total);
  for (int p = 0; p < fmt->numPlanes; p++) {
    fr->wptr[p] = fr->owned + offs[p];
    fr->rptr[p] = fr->wptr[p];
  }
  return fr;
}

VSFrameRef* VS_CC fake_new_frame(const VSFormat* fmt, int w, int h, const VSFrameRef*, VSCore*) noexcept {
  return alloc_frame(fmt, w, h);
}

// Plane copies from planeSrc[0] (the only use, pshow's Create(true), passes
// a one-element array whose frame holds every plane).
VSFrameRef* VS_CC fake_new_frame2(const VSFormat* fmt, int w, int h, const VSFrameRef** planeSrc,
                                  const int*, const VSFrameRef*, VSCore*) noexcept {
  VSFrameRef* fr = alloc_frame(fmt, w, h);
  const VSFrameRef* s = planeSrc ? planeSrc[0] : nullptr;
  if (s)
    for (int p = 0; p < fmt->numPlanes; p++)
      for (int y = 0; y < fr->height[p]; y++)
        memcpy(fr->wptr[p] + size_t(y) * fr->stride[p], s->rptr[p] + size_t(y) * s->stride[p],
               size_t(fr->width[p]) * fmt->bytesPerSample);
  return fr;
}

int VS_CC fake_stride(const VSFrameRef* f, int p) noexcept { return f->stride[p]; }
const uint8_t* VS_CC fake_rptr(const VSFrameRef* f, int p) noexcept { return f->rptr[p]; }
uint8_t* VS_CC fake_wptr(VSFrameRef* f, int p) noexcept { return f->wptr[p]; }
const VSFormat* VS_CC fake_format(const VSFrameRef* f) noexcept { return f->fmt; }
int VS_CC fake_width(const VSFrameRef* f, int p) noexcept { return f->width[p]; }
int VS_CC fake_height(const VSFrameRef* f, int p) noexcept { return f->height[p]; }

const VSAPI* fake_api() {
  static VSAPI api = [] {
    VSAPI a;
    memset(&a, 0, sizeof(a));
    a.cloneFrameRef = fake_clone;
    a.freeFrame = fake_free;
    a.newVideoFrame = fake_new_frame;
    a.newVideoFrame2 = fake_new_frame2;
    a.getStride = fake_stride;
    a.getReadPtr = fake_rptr;
    a.getWritePtr = fake_wptr;
    a.getFrameFormat = fake_format;
    a.getFrameWidth = fake_width;
    a.getFrameHeight = fake_height;
    return a;
  }();
  return &api;
}

struct MapDelegator final : InDelegator {
  std::map<std::string, std::vector<std::string>> kv;
  explicit MapDelegator(const char* args) {
    std::stringstream ss(args ? args : "");
    std::string item;
    while (std::getline(ss, item, ';')) {
      auto eq = item.find('=');
      if (eq == std::string::npos) continue;
      std::vector<std::string> vals;
      std::stringstream vs(item.substr(eq + 1));
      std::string v;
      while (std::getline(vs, v, ',')) vals.push_back(v);
      kv[item.substr(0, eq)] = vals;
    }
  }
  const std::string* get(const char* n) {
    auto it = kv.find(n);
    return (it == kv.end() || it->second.empty()) ? nullptr : &it->second[0];
  }
  void Read(const char* n, int& o) override { if (auto s = get(n)) o = std::stoi(*s); }
  void Read(const char* n, int64_t& o) override { if (auto s = get(n)) o = std::stoll(*s); }
  void Read(const char* n, float& o) override { if (auto s = get(n)) o = static_cast<float>(std::stod(*s)); }
  void Read(const char* n, double& o) override { if (auto s = get(n)) o = std::stod(*s); }
  void Read(const char* n, bool& o) override { if (auto s = get(n)) o = std::stoll(*s) != 0; }
  void Read(const char* n, std::string& o) override { if (auto s = get(n)) o = *s; }
  void Read(const char* n, std::vector<int>& o) override {
    auto it = kv.find(n);
    if (it == kv.end()) return;
    o.clear();
    for (auto& v : it->second) o.push_back(std::stoi(v));
  }
  void Read(const char*, std::vector<int64_t>&) override {}
  void Read(const char*, std::vector<float>&) override {}
  void Read(const char*, std::vector<double>&) override {}
  void Read(const char*, std::vector<bool>&) override {}
  void Read(const char*, void*& o) override { o = nullptr; }
  void Free(void*& c) override { c = nullptr; }
};

struct RefClip;

struct StoreFetch final : FetchFrameFunctor {
  RefClip* clip;
  explicit StoreFetch(RefClip* c) : clip(c) {}
  DSFrame operator()(int n) override;
};

struct RefClip {
  VSFormat fmt{};
  int width{0}, height{0}, frames{0};
  std::vector<VSFrameRef*> src;  // caller memory, external
  StoreFetch fetch{this};
  FFT3D filter;
  ~RefClip() {
    for (auto* f : src) delete f;
  }
};

DSFrame StoreFetch::operator()(int n) {
  n = std::max(0, std::min(n, clip->frames - 1));
  return DSFrame(fake_clone(clip->src[n]), nullptr, fake_api());
}

void set_err(char* err, int len, const char* prefix, const char* msg) {
  if (err && len > 0) snprintf(err, size_t(len), "%s%s", prefix, msg);
}

}  // namespace

extern "C" {

// family: 0 gray, 1 yuv, 2 rgb; sample_type: 0 integer, 1 float.
// planes[f * nplanes + p] / strides[p] describe the in-memory clip (borrowed
// for the lifetime of the handle).
void* nf3dref_create(const char* args, int family, int sample_type, int bits, int ssw, int ssh,
                     int width, int height, int nframes, const void* const* planes,
                     const int* strides, char* err, int errlen) {
  auto clip = std::make_unique<RefClip>();
  VSFormat& f = clip->fmt;
  f.colorFamily = family == 0 ? cmGray : family == 1 ? cmYUV : cmRGB;
  f.sampleType = sample_type ? stFloat : stInteger;
  f.bitsPerSample = bits;
  f.bytesPerSample = bits <= 8 ? 1 : bits <= 16 ? 2 : 4;
  f.subSamplingW = family == 1 ? ssw : 0;
  f.subSamplingH = family == 1 ? ssh : 0;
  f.numPlanes = family == 0 ? 1 : 3;
  snprintf(f.name, sizeof(f.name), "fake");
  clip->width = width;
  clip->height = height;
  clip->frames = nframes;
  for (int n = 0; n < nframes; n++) {
    auto* fr = new VSFrameRef();
    fr->fmt = &clip->fmt;
    fr->external = true;
    for (int p = 0; p < f.numPlanes; p++) {
      plane_dims(&f, width, height, p, fr->width[p], fr->height[p]);
      fr->stride[p] = strides[p];
      fr->rptr[p] = static_cast<const uint8_t*>(planes[size_t(n) * f.numPlanes + p]);
    }
    clip->src.push_back(fr);
  }
  try {
    MapDelegator in(args);
    DSVideoInfo vi(DSFormat(&clip->fmt), 25, 1, width, height, nframes);
    clip->filter.Initialize(&in, vi, &clip->fetch);
  } catch (const char* e) {
    set_err(err, errlen, "FFT3D: ", e);
    return nullptr;
  } catch (const std::exception& e) {
    set_err(err, errlen, "FFT3D: ", e.what());
    return nullptr;
  }
  return clip.release();
}

// Runs the front end's request window (src/fft3d.hpp:201-219) and GetFrame
// (src/fft3d.hpp:221-264), then copies every plane of the result out.
int nf3dref_get_frame(void* h, int n, void* const* dst, const int* dst_strides, char* err, int errlen) {
  auto* clip = static_cast<RefClip*>(h);
  try {
    std::unordered_map<int, DSFrame> in_frames;
    for (int i : clip->filter.RequestReferenceFrames(n))
      in_frames[i] = DSFrame(fake_clone(clip->src[i]), nullptr, fake_api());
    DSFrame out = clip->filter.GetFrame(n, in_frames);
    const VSFrameRef* fr = out.ToVSFrame();
    for (int p = 0; p < clip->fmt.numPlanes; p++) {
      size_t row = size_t(fr->width[p]) * clip->fmt.bytesPerSample;
      for (int y = 0; y < fr->height[p]; y++)
        memcpy(static_cast<uint8_t*>(dst[p]) + size_t(y) * dst_strides[p],
               fr->rptr[p] + size_t(y) * fr->stride[p], row);
    }
    fake_free(fr);
  } catch (const char* e) {
    set_err(err, errlen, "FFT3D: ", e);
    return -1;
  }
  return 0;
}

void nf3dref_free(void* h) { delete static_cast<RefClip*>(h); }

// Frames [first, first+count) through the front end from `threads` host
// threads pulling frame numbers from a shared counter (a VapourSynth core
// issuing concurrent getFrame calls to an fmParallel filter); outputs go to
// per-thread scratch frames.  Returns 0, or -1 with the first error.
int nf3dref_run_parallel(void* h, int first, int count, int threads, char* err, int errlen) {
  auto* clip = static_cast<RefClip*>(h);
  std::atomic<int> next{first};
  std::atomic<int> failed{0};
  std::mutex err_mu;
  auto worker = [&]() {
    VSFrameRef* scratch = alloc_frame(&clip->fmt, clip->width, clip->height);
    void* dst[4];
    int ds[4];
    for (int p = 0; p < clip->fmt.numPlanes; p++) {
      dst[p] = scratch->wptr[p];
      ds[p] = scratch->stride[p];
    }
    char local[256];
    for (;;) {
      int n = next.fetch_add(1);
      if (n >= first + count || failed.load()) break;
      if (nf3dref_get_frame(h, n, dst, ds, local, sizeof(local)) != 0) {
        std::lock_guard<std::mutex> lock(err_mu);
        if (!failed.exchange(1)) set_err(err, errlen, "", local);
      }
    }
    fake_free(scratch);
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < std::max(1, threads); t++) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  return failed.load() ? -1 : 0;
}

const char* nf3dref_vs_params() {
  static std::string s = FFT3D().VSParams();
  return s.c_str();
}

const char* nf3dref_avs_params() {
  static std::string s = FFT3D().AVSParams();
  return s.c_str();
}

}  // extern "C"
