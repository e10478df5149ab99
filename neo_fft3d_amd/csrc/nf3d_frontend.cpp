// The FFT3D filter front end over the per-plane GPU engines: nf3d_filter_*
// and the clip-level nf3d_clip_* of the C ABI.
//
// Mirrors src/fft3d.hpp:12-364: the 41-parameter surface with its defaults
// (:77-139), crop validation (:141-147), VapourSynth `planes` / AviSynth+
// y,u,v plane modes (:149-169: 3 process, 2 copy, 1 leave), one engine per
// processed plane (:182-198), the frame-request window (:201-219) and the
// border / unprocessed-plane copies of GetFrame (:221-352).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/nf3d.h"

namespace {

enum PType { PClip, PInt, PFloat, PBool };
struct PDesc {
  const char* name;
  PType type;
  bool array;
  bool avs;
  bool vs;
};

// FFT3D::Params() order and types (src/fft3d.hpp:29-73).
const PDesc kParams[] = {
    {"clip", PClip, false, true, true},      {"sigma", PFloat, false, true, true},
    {"beta", PFloat, false, true, true},     {"planes", PInt, true, false, true},
    {"bw", PInt, false, true, true},         {"bh", PInt, false, true, true},
    {"bt", PInt, false, true, true},         {"ow", PInt, false, true, true},
    {"oh", PInt, false, true, true},         {"kratio", PFloat, false, true, true},
    {"sharpen", PFloat, false, true, true},  {"scutoff", PFloat, false, true, true},
    {"svr", PFloat, false, true, true},      {"smin", PFloat, false, true, true},
    {"smax", PFloat, false, true, true},     {"measure", PBool, false, true, true},
    {"interlaced", PBool, false, true, true}, {"wintype", PInt, false, true, true},
    {"pframe", PInt, false, true, true},     {"px", PInt, false, true, true},
    {"py", PInt, false, true, true},         {"pshow", PBool, false, true, true},
    {"pcutoff", PFloat, false, true, true},  {"pfactor", PFloat, false, true, true},
    {"sigma2", PFloat, false, true, true},   {"sigma3", PFloat, false, true, true},
    {"sigma4", PFloat, false, true, true},   {"degrid", PFloat, false, true, true},
    {"dehalo", PFloat, false, true, true},   {"hr", PFloat, false, true, true},
    {"ht", PFloat, false, true, true},       {"y", PInt, false, true, false},
    {"u", PInt, false, true, false},         {"v", PInt, false, true, false},
    {"l", PInt, false, true, true},          {"t", PInt, false, true, true},
    {"r", PInt, false, true, true},          {"b", PInt, false, true, true},
    {"opt", PInt, false, true, true},        {"ncpu", PInt, false, true, true},
    {"mt", PBool, false, true, true},
};

std::string make_vs_signature() {
  // include/dualsynth/ds_filter.hpp:31-54 (clip is required, the rest optional)
  std::string s;
  for (const auto& p : kParams) {
    if (!p.vs) continue;
    const char* t = p.type == PClip ? "clip" : p.type == PFloat ? "float" : "int";
    s += std::string(p.name) + ":" + t + (p.array ? "[]" : "") + (p.type == PClip ? "" : ":opt") + ";";
  }
  return s;
}

std::string make_avs_signature() {
  // include/dualsynth/ds_filter.hpp:55-69
  std::string s;
  for (const auto& p : kParams) {
    if (!p.avs) continue;
    char t = p.type == PClip ? 'c' : p.type == PInt ? 'i' : p.type == PFloat ? 'f' : 'b';
    if (p.type != PClip) s += std::string("[") + p.name + "]";
    s += t;
  }
  return s;
}

struct Args {
  std::map<std::string, std::vector<std::string>> kv;
  explicit Args(const char* a) {
    std::stringstream ss(a ? a : "");
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
  bool has(const char* n) const { return kv.count(n) > 0; }
  void read(const char* n, int& o) const {
    auto it = kv.find(n);
    if (it != kv.end() && !it->second.empty()) o = std::atoi(it->second[0].c_str());
  }
  void read(const char* n, float& o) const {
    auto it = kv.find(n);
    if (it != kv.end() && !it->second.empty()) o = (float)std::atof(it->second[0].c_str());
  }
  void read_list(const char* n, std::vector<int>& o) const {
    auto it = kv.find(n);
    if (it == kv.end()) return;
    o.clear();
    for (auto& v : it->second) o.push_back(std::atoi(v.c_str()));
  }
};

void put_err(char* err, size_t len, const std::string& m) {
  if (err && len) {
    size_t n = std::min(len - 1, m.size());
    memcpy(err, m.data(), n);
    err[n] = 0;
  }
}

}  // namespace

struct nf3d_filter {
  nf3d_params p{};
  nf3d_format fmt{};
  int width = 0, height = 0, frames = 0;
  int process[4] = {2, 2, 2, 2};
  nf3d_engine* eng[4] = {nullptr, nullptr, nullptr, nullptr};
  int engine_count = 0, copy_count = 0;
  bool crop = false;
  bool pattern_pending = false;
  int bps = 1;

  ~nf3d_filter() {
    for (auto* e : eng)
      if (e) nf3d_engine_destroy(e);
  }

  void plane_dims(int pl, int& w, int& h) const {
    bool sub = fmt.color_family == 1 && pl > 0 && pl < 3;
    w = sub ? width >> fmt.ssw : width;
    h = sub ? height >> fmt.ssh : height;
  }
};

extern "C" {

const char* nf3d_vs_signature(void) {
  static const std::string s = make_vs_signature();
  return s.c_str();
}

const char* nf3d_avs_signature(void) {
  static const std::string s = make_avs_signature();
  return s.c_str();
}

const char* nf3d_version(void) { return "Neo FFT3D Filter (MI355X/gfx950) r1"; }

int nf3d_filter_create(const char* args, const nf3d_format* fmt, int width, int height, int frames,
                       nf3d_filter** out, char* err, size_t errlen) {
  Args in(args);
  auto* f = new nf3d_filter();
  f->fmt = *fmt;
  f->width = width;
  f->height = height;
  f->frames = frames;
  f->bps = fmt->is_float ? 4 : fmt->bits_per_sample <= 8 ? 1 : 2;
  float sigma1 = 2.0f;
  in.read("sigma", sigma1);
  nf3d_params& p = f->p;
  nf3d_params_defaults(&p, sigma1);
  in.read("beta", p.beta);
  in.read("bw", p.bw);
  in.read("bh", p.bh);
  in.read("bt", p.bt);
  in.read("ow", p.ow);
  in.read("oh", p.oh);
  in.read("kratio", p.kratio);
  in.read("sharpen", p.sharpen);
  in.read("scutoff", p.scutoff);
  in.read("svr", p.svr);
  in.read("smin", p.smin);
  in.read("smax", p.smax);
  in.read("measure", p.measure);
  in.read("interlaced", p.interlaced);
  in.read("wintype", p.wintype);
  in.read("pframe", p.pframe);
  in.read("px", p.px);
  in.read("py", p.py);
  in.read("pshow", p.pshow);
  in.read("pcutoff", p.pcutoff);
  in.read("pfactor", p.pfactor);
  in.read("sigma2", p.sigma2);
  in.read("sigma3", p.sigma3);
  in.read("sigma4", p.sigma4);
  in.read("degrid", p.degrid);
  in.read("dehalo", p.dehalo);
  in.read("hr", p.hr);
  in.read("ht", p.ht);
  in.read("l", p.l);
  p.l = std::max(p.l, 0);
  in.read("t", p.t);
  p.t = std::max(p.t, 0);
  in.read("r", p.r);
  p.r = std::max(p.r, 0);
  in.read("b", p.b);
  p.b = std::max(p.b, 0);
  in.read("opt", p.opt);
  p.measure = p.measure != 0;
  p.interlaced = p.interlaced != 0;
  p.pshow = p.pshow != 0;
  f->crop = p.l > 0 || p.r > 0 || p.t > 0 || p.b > 0;
  if (p.l + p.r >= width) {
    put_err(err, errlen, "Width cannot be cropped to zero or below");
    delete f;
    return -1;
  }
  if (p.t + p.b >= height) {
    put_err(err, errlen, "Height cannot be cropped to zero or below");
    delete f;
    return -1;
  }
  if (in.has("planes") || !(in.has("y") || in.has("u") || in.has("v"))) {
    std::vector<int> user{0, 1, 2};
    in.read_list("planes", user);
    for (int i = 0; i < 4; i++) f->process[i] = 2;
    for (int q : user)
      if (q >= 0 && q < fmt->num_planes) f->process[q] = 3;
  } else {
    f->process[0] = f->process[1] = f->process[2] = 3;
    in.read("y", f->process[0]);
    in.read("u", f->process[1]);
    in.read("v", f->process[2]);
  }
  if (fmt->num_planes < 1 || fmt->num_planes > 4 || (fmt->color_family == 0 && fmt->num_planes != 1)) {
    put_err(err, errlen, "formats have 1 (gray), 3 or 4 (with alpha) planes");
    delete f;
    return -1;
  }
  // the alpha plane of a YUV format is filtered with the chroma geometry
  // (src/fft3d.hpp:186-191), which for a subsampled format covers part of
  // the plane while the whole plane is copied out (:292-313): the reference
  // leaves the rest undefined, so that case is refused
  if (fmt->num_planes == 4 && f->process[3] == 3 && fmt->color_family == 1 && (fmt->ssw || fmt->ssh)) {
    put_err(err, errlen, "the alpha plane of a subsampled YUVA clip cannot be filtered");
    delete f;
    return -1;
  }
  for (int i = 0; i < fmt->num_planes; i++) {
    if (f->process[i] == 3) {
      std::string e(512, '\0');
      if (nf3d_engine_create(&p, fmt, width, height, frames, i, &f->eng[i], &e[0], e.size()) != 0) {
        put_err(err, errlen, e.c_str());
        delete f;
        return -1;
      }
      f->engine_count++;
      // "dft=1": the matrix-core DFT form (an extension; not part of the
      // reference's parameter list or signatures)
      int dft = 0;
      in.read("dft", dft);
      if (dft && nf3d_engine_set_dft(f->eng[i], dft, &e[0], e.size()) != 0) {
        put_err(err, errlen, e.c_str());
        delete f;
        return -1;
      }
      // "pregion=rx[,ry]": per-region noise patterns (an extension)
      std::vector<int> preg;
      in.read_list("pregion", preg);
      if (!preg.empty()) {
        const int rx = preg[0], ry = preg.size() > 1 ? preg[1] : preg[0];
        if (nf3d_engine_set_pattern_regions(f->eng[i], rx, ry, &e[0], e.size()) != 0) {
          put_err(err, errlen, e.c_str());
          delete f;
          return -1;
        }
      }
    } else if (f->process[i] == 2) {
      f->copy_count++;
    }
  }
  // the estimated pattern needs frame pframe (src/fft3d_engine.h:485-504);
  // a pattern built from sigma2..4 does not (src/fft3d_engine.h:366-371)
  bool sigmas = p.sigma2 != p.sigma || p.sigma3 != p.sigma || p.sigma4 != p.sigma;
  f->pattern_pending = f->engine_count > 0 && p.pfactor != 0 && !(sigmas && p.pfactor == 0);
  *out = f;
  return 0;
}

int nf3d_filter_request_frames(nf3d_filter* f, int n, int* out, int cap) {
  std::vector<int> req;
  if (f->p.bt > 1 && f->engine_count > 0) {
    int from = std::max(n - f->p.bt / 2, 0);
    int to = std::min(n + (f->p.bt - 1) / 2, f->frames - 1);
    for (int i = from; i <= to; i++) {
      for (int j = 0; j < f->fmt.num_planes; j++)
        if (f->process[j] == 3) nf3d_engine_cache_refresh(f->eng[j], i);
      req.push_back(i);
    }
  } else {
    req.push_back(n);
  }
  if (f->pattern_pending) {
    int pf = std::max(0, std::min(f->p.pframe, f->frames - 1));
    if (std::find(req.begin(), req.end(), pf) == req.end()) req.push_back(pf);
  }
  int k = 0;
  for (int v : req)
    if (k < cap) out[k++] = v;
  return k;
}

int nf3d_filter_mt_mode(const nf3d_filter* f) { return f->p.bt == 0 ? 0 : 1; }

int nf3d_filter_get_frame(nf3d_filter* f, int n, const nf3d_frame_planes* frames, int nframes,
                          void* const* dst, const int* dst_stride, char* err, size_t errlen) {
  const nf3d_frame_planes* cur = nullptr;
  for (int i = 0; i < nframes; i++)
    if (frames[i].n == n) cur = &frames[i];
  if (!cur) {
    put_err(err, errlen, "frame " + std::to_string(n) + " was not provided");
    return -1;
  }
  std::vector<nf3d_frame_ref> refs(nframes);
  for (int pl = 0; pl < f->fmt.num_planes; pl++) {
    int w, h;
    f->plane_dims(pl, w, h);
    auto copy_src = [&]() {
      const uint8_t* s = static_cast<const uint8_t*>(cur->data[pl]);
      uint8_t* d = static_cast<uint8_t*>(dst[pl]);
      for (int y = 0; y < h; y++)
        memcpy(d + (size_t)y * dst_stride[pl], s + (size_t)y * cur->stride[pl], (size_t)w * f->bps);
    };
    if (f->process[pl] == 3) {
      if (f->crop) copy_src();  // borders keep the source (src/fft3d.hpp:315-351)
      for (int i = 0; i < nframes; i++) refs[i] = nf3d_frame_ref{frames[i].n, frames[i].data[pl], frames[i].stride[pl]};
      if (nf3d_engine_get_frame(f->eng[pl], n, refs.data(), nframes, dst[pl], dst_stride[pl], err, errlen) != 0)
        return -1;
    } else if (f->process[pl] == 2) {
      copy_src();
    }
  }
  f->pattern_pending = false;
  return 0;
}

void nf3d_filter_destroy(nf3d_filter* f) { delete f; }

// ----------------------------------------------------------- clip-level
struct nf3d_clip {
  nf3d_filter* f = nullptr;
  int nplanes = 1;
  int frames = 0;
  std::vector<const void*> planes;
  std::vector<int> strides;
  std::mutex mu;
};

void* nf3d_clip_create(const char* args, int family, int sample_type, int bits, int ssw, int ssh, int width,
                       int height, int nframes, const void* const* planes, const int* strides, char* err,
                       int errlen) {
  nf3d_format fmt{};
  // families 3 / 4: YUVA / RGBA (a fourth, alpha plane at the luma size)
  const int alpha = family == 3 || family == 4;
  if (family == 3) family = 1;
  if (family == 4) family = 2;
  fmt.color_family = family;
  fmt.is_float = sample_type != 0;
  fmt.bits_per_sample = bits;
  fmt.ssw = family == 1 ? ssw : 0;
  fmt.ssh = family == 1 ? ssh : 0;
  fmt.num_planes = family == 0 ? 1 : 3 + alpha;
  auto* c = new nf3d_clip();
  c->nplanes = fmt.num_planes;
  c->frames = nframes;
  c->planes.assign(planes, planes + (size_t)nframes * fmt.num_planes);
  c->strides.assign(strides, strides + fmt.num_planes);
  std::string e(512, '\0');
  if (nf3d_filter_create(args, &fmt, width, height, nframes, &c->f, &e[0], e.size()) != 0) {
    std::string m = std::string("FFT3D: ") + e.c_str();
    put_err(err, (size_t)errlen, m);
    delete c;
    return nullptr;
  }
  return c;
}

int nf3d_clip_get_frame(void* h, int n, void* const* dst, const int* dst_strides, char* err, int errlen) {
  auto* c = static_cast<nf3d_clip*>(h);
  std::lock_guard<std::mutex> lock(c->mu);
  int req[8];
  int k = nf3d_filter_request_frames(c->f, n, req, 8);
  std::vector<nf3d_frame_planes> fr(k);
  for (int i = 0; i < k; i++) {
    fr[i].n = req[i];
    for (int p = 0; p < c->nplanes; p++) {
      fr[i].data[p] = c->planes[(size_t)req[i] * c->nplanes + p];
      fr[i].stride[p] = c->strides[p];
    }
  }
  std::string e(512, '\0');
  if (nf3d_filter_get_frame(c->f, n, fr.data(), k, dst, dst_strides, &e[0], e.size()) != 0) {
    put_err(err, (size_t)errlen, std::string("FFT3D: ") + e.c_str());
    return -1;
  }
  return 0;
}

void nf3d_clip_free(void* h) {
  auto* c = static_cast<nf3d_clip*>(h);
  if (!c) return;
  nf3d_filter_destroy(c->f);
  delete c;
}

}  // extern "C"
