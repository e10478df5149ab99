// Seeded synthetic clips for the benchmark workloads and parity fixtures.
//
// Counter-based randomness (splitmix64 finaliser over a key built from
// seed/frame/plane/pixel, Box-Muller for Gaussians) so any frame can be
// generated independently, in parallel, and identically on every machine.
//
//   nf3d_synth_gratings  gradient + 3 drifting sinusoidal gratings (period
//                        8-64 px, 1 px/frame) + N(0, noise^2), one plane
//   nf3d_synth_rects_*   gradient + moving fBm-textured rectangles (size
//                        32-400 px, velocity +-4 px/frame, texture amplitude
//                        40) + N(0, noise^2) per plane, YUV 4:2:0 / 4:4:4
//
// Content and noise levels are in 8-bit units.  Integer output is scaled by
// 2^(bits-8), rounded and clipped to [0, 2^bits-1] (uint8 for 8 bits, uint16
// otherwise); float output is scaled by 1/255 (value range [0,1]).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline double u01(uint64_t key) { return (double)(mix64(key) >> 11) * (1.0 / 9007199254740992.0); }

inline double gauss(uint64_t key) {
  double u1 = ((double)(mix64(2 * key) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  double u2 = u01(2 * key + 1);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

inline uint64_t key3(uint64_t seed, uint64_t a, uint64_t b) {
  return mix64(mix64(seed ^ 0xA5A5A5A5ull) + 0x1000003ull * a) + b;
}

template <typename T>
void store(void* dst, int stride, int x, int y, double v, int bits) {
  T* row = reinterpret_cast<T*>(static_cast<uint8_t*>(dst) + (size_t)y * stride);
  if constexpr (sizeof(T) == 4) {
    row[x] = (float)(v / 255.0);
  } else {
    double m = (double)((1 << bits) - 1);
    double r = std::nearbyint(v * (double)(1 << (bits - 8)));
    row[x] = (T)(r < 0 ? 0 : r > m ? m : r);
  }
}

void put(void* dst, int stride, int x, int y, double v, int bits, int is_float) {
  if (is_float) store<float>(dst, stride, x, y, v, bits);
  else if (bits <= 8) store<uint8_t>(dst, stride, x, y, v, bits);
  else store<uint16_t>(dst, stride, x, y, v, bits);
}

// Value noise on an integer lattice, smoothstep interpolation, range [-1,1].
double value_noise(uint64_t seed, double u, double v) {
  double fu = std::floor(u), fv = std::floor(v);
  int64_t iu = (int64_t)fu, iv = (int64_t)fv;
  double tu = u - fu, tv = v - fv;
  tu = tu * tu * (3 - 2 * tu);
  tv = tv * tv * (3 - 2 * tv);
  auto L = [&](int64_t a, int64_t b) { return 2 * u01(key3(seed, (uint64_t)a, (uint64_t)b)) - 1; };
  double a = L(iu, iv), b = L(iu + 1, iv), c = L(iu, iv + 1), d = L(iu + 1, iv + 1);
  return (a * (1 - tu) + b * tu) * (1 - tv) + (c * (1 - tu) + d * tu) * tv;
}

double fbm(uint64_t seed, double u, double v) {
  double s = 0, amp = 1, norm = 0, f = 1.0 / 24.0;
  for (int o = 0; o < 5; o++) {
    s += amp * value_noise(seed + 977 * o, u * f, v * f);
    norm += amp;
    amp *= 0.5;
    f *= 2;
  }
  return s / norm;
}

struct Rect {
  double x0, y0, vx, vy;
  int w, h;
  double level, cu, cv;
  std::vector<float> tex;  // w*h, amplitude-scaled
};

}  // namespace

extern "C" {

// One plane of frame `frame`: gradient + 3 drifting gratings + noise.
void nf3d_synth_gratings(uint64_t seed, int frame, int width, int height, int bits, int is_float,
                         double noise, void* dst, int stride) {
  double per[3], ang[3], ph[3];
  for (int k = 0; k < 3; k++) {
    per[k] = 8 + 56 * u01(key3(seed, 1, k));
    ang[k] = 3.141592653589793 * u01(key3(seed, 2, k));
    ph[k] = 6.283185307179586 * u01(key3(seed, 3, k));
  }
  uint64_t nk = key3(seed, 100 + (uint64_t)frame, 0);
  for (int y = 0; y < height; y++)
    for (int x = 0; x < width; x++) {
      double v = 64 + 64.0 * x / (width > 1 ? width - 1 : 1) + 48.0 * y / (height > 1 ? height - 1 : 1);
      for (int k = 0; k < 3; k++) {
        double d = x * std::cos(ang[k]) + y * std::sin(ang[k]) - frame;
        v += 16 * std::sin(6.283185307179586 * d / per[k] + ph[k]);
      }
      v += noise * gauss(mix64(nk) + (uint64_t)y * (uint64_t)width + (uint64_t)x);
      put(dst, stride, x, y, v, bits, is_float);
    }
}

// Moving-rectangles clip: the handle holds the rectangles and their textures
// (computed once), frames are then generated independently.
void* nf3d_synth_rects_create(uint64_t seed, int width, int height) {
  const int nrect = 24;
  auto* rs = new std::vector<Rect>(nrect);
  for (int k = 0; k < nrect; k++) {
    Rect& r = (*rs)[k];
    r.w = 32 + (int)(369 * u01(key3(seed, 10, k)));
    r.h = 32 + (int)(369 * u01(key3(seed, 11, k)));
    r.x0 = u01(key3(seed, 12, k)) * width - r.w / 2.0;
    r.y0 = u01(key3(seed, 13, k)) * height - r.h / 2.0;
    r.vx = -4 + 8 * u01(key3(seed, 14, k));
    r.vy = -4 + 8 * u01(key3(seed, 15, k));
    r.level = 40 + 170 * u01(key3(seed, 16, k));
    r.cu = 80 + 96 * u01(key3(seed, 17, k));
    r.cv = 80 + 96 * u01(key3(seed, 18, k));
    r.tex.resize((size_t)r.w * r.h);
    for (int v = 0; v < r.h; v++)
      for (int u = 0; u < r.w; u++) r.tex[(size_t)v * r.w + u] = (float)fbm(seed + 31 * k, u, v);
  }
  return rs;
}

void nf3d_synth_rects_free(void* h) { delete static_cast<std::vector<Rect>*>(h); }

// All planes of frame `frame`.  ssw/ssh: chroma subsampling shifts (1,1 for
// 4:2:0; 0,0 for 4:4:4).  dst[0..2]/strides[0..2]: the three planes.
void nf3d_synth_rects_frame(void* h, uint64_t seed, int frame, int width, int height, int ssw, int ssh,
                            int bits, int is_float, double noise_luma, double noise_chroma,
                            void* const* dst, const int* strides) {
  const std::vector<Rect>& rs = *static_cast<std::vector<Rect>*>(h);
  for (int p = 0; p < 3; p++) {
    int pw = p ? width >> ssw : width, ph = p ? height >> ssh : height;
    int sx = p ? ssw : 0, sy = p ? ssh : 0;
    std::vector<double> img((size_t)pw * ph);
    for (int y = 0; y < ph; y++)
      for (int x = 0; x < pw; x++) {
        double gx = (double)(x << sx) / width, gy = (double)(y << sy) / height;
        img[(size_t)y * pw + x] = p == 0 ? 40 + 100 * gx + 40 * gy
                                         : 128 + (p == 1 ? 24 : -24) * (gx - 0.5) + 12 * (gy - 0.5);
      }
    for (const Rect& r : rs) {
      int rx = (int)std::floor(r.x0 + r.vx * frame), ry = (int)std::floor(r.y0 + r.vy * frame);
      int x0 = rx >> sx, y0 = ry >> sy, x1 = (rx + r.w) >> sx, y1 = (ry + r.h) >> sy;
      double lvl = p == 0 ? r.level : p == 1 ? r.cu : r.cv;
      double amp = p == 0 ? 40.0 : 20.0;
      for (int y = y0 < 0 ? 0 : y0; y < (y1 > ph ? ph : y1); y++)
        for (int x = x0 < 0 ? 0 : x0; x < (x1 > pw ? pw : x1); x++) {
          int u = (x << sx) - rx, v = (y << sy) - ry;
          u = u < 0 ? 0 : u >= r.w ? r.w - 1 : u;
          v = v < 0 ? 0 : v >= r.h ? r.h - 1 : v;
          img[(size_t)y * pw + x] = lvl + amp * r.tex[(size_t)v * r.w + u];
        }
    }
    double nz = p == 0 ? noise_luma : noise_chroma;
    uint64_t nk = key3(seed, 200 + (uint64_t)frame, (uint64_t)p);
    for (int y = 0; y < ph; y++)
      for (int x = 0; x < pw; x++) {
        double v = img[(size_t)y * pw + x] + nz * gauss(mix64(nk) + (uint64_t)y * (uint64_t)pw + (uint64_t)x);
        put(dst[p], strides[p], x, y, v, bits, is_float);
      }
  }
}

}  // extern "C"
