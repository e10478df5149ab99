// Register-resident complex FFTs of compile-time length for gfx950.
//
// One lane owns one whole transform (a block row pair or a block column), so
// a length-N transform is N float2 values in VGPRs, fully unrolled: every
// index and every twiddle is a compile-time constant, and the recursion below
// compiles to straight-line butterflies with no LDS traffic and no shuffles.
//
// Decimation in time, mixed radix (4, 2, 3, 5; any other prime factor falls
// back to a direct DFT of that length).  Twiddles are evaluated in double at
// compile time (constexpr range-reduced Taylor series) and rounded once to
// float, so they are correctly rounded like FFTW's precomputed tables.
// Unnormalised, FFTW sign convention: SIGN = -1 forward, +1 inverse.
#pragma once

#include <hip/hip_runtime.h>
#include <type_traits>
#include <utility>

namespace nf3d {

// ---------------------------------------------------------------- constexpr trig
namespace ct {
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x) {  // |x| <= pi/4
  double x2 = x * x, term = x, sum = x;
  for (int k = 1; k < 14; k++) {
    term *= -x2 / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {  // |x| <= pi/4
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; k++) {
    term *= -x2 / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
// cos and sin of 2*pi*a/n, with the octant folded exactly in integers.
constexpr double cos2pi(long a, long n) {
  a %= n;
  if (a < 0) a += n;
  // work with m = 8a; octant o = floor(8a/n)
  long m8 = 8 * a;
  long o = m8 / n;
  double r = 2.0 * kPi * (double)(m8 - o * n) / (8.0 * (double)n);  // in [0, pi/4)
  switch (o) {
    case 0: return taylor_cos(r);
    case 1: return taylor_sin(kPi / 4 - r);
    case 2: return -taylor_sin(r);
    case 3: return -taylor_cos(kPi / 4 - r);
    case 4: return -taylor_cos(r);
    case 5: return -taylor_sin(kPi / 4 - r);
    case 6: return taylor_sin(r);
    default: return taylor_cos(kPi / 4 - r);
  }
}
constexpr double sin2pi(long a, long n) {
  a %= n;
  if (a < 0) a += n;
  long m8 = 8 * a;
  long o = m8 / n;
  double r = 2.0 * kPi * (double)(m8 - o * n) / (8.0 * (double)n);
  switch (o) {
    case 0: return taylor_sin(r);
    case 1: return taylor_cos(kPi / 4 - r);
    case 2: return taylor_cos(r);
    case 3: return taylor_sin(kPi / 4 - r);
    case 4: return -taylor_sin(r);
    case 5: return -taylor_cos(kPi / 4 - r);
    case 6: return -taylor_cos(r);
    default: return -taylor_sin(kPi / 4 - r);
  }
}
}  // namespace ct

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// The library is compiled with -ffp-contract=off (the spectral pass and the
// overlap-add follow the reference's float rounding); the FFT opts into
// fused multiply-adds explicitly.
__device__ __forceinline__ float2 cmul(float2 a, float wr, float wi) {
  return make_float2(__builtin_fmaf(a.x, wr, -a.y * wi), __builtin_fmaf(a.x, wi, a.y * wr));
}
__device__ __forceinline__ float2 caxpy(float2 y, float s, float2 x) {  // y + s*x
  return make_float2(__builtin_fmaf(s, x.x, y.x), __builtin_fmaf(s, x.y, y.y));
}
// (SIGN * i) * a
template <int SIGN>
__device__ __forceinline__ float2 mul_i(float2 a) {
  return SIGN < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

constexpr int pick_radix(int n) {
  if (n % 4 == 0 && n != 8) return 4;  // 8 = 2*4 keeps the 4-point core innermost
  if (n % 2 == 0) return 2;
  if (n % 3 == 0) return 3;
  if (n % 5 == 0) return 5;
  for (int p = 7; p * p <= n; p += 2)
    if (n % p == 0) return p;
  return n;
}

// Radix-R butterfly, in place on t[0..R-1]: t[q] <- sum_r t[r] w_R^(SIGN*r*q).
template <int R, int SIGN>
__device__ __forceinline__ void butterfly(float2 (&t)[R]) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    float2 a = t[0], b = t[1];
    t[0] = cadd(a, b);
    t[1] = csub(a, b);
  } else if constexpr (R == 4) {
    float2 s02 = cadd(t[0], t[2]), d02 = csub(t[0], t[2]);
    float2 s13 = cadd(t[1], t[3]), d13 = mul_i<SIGN>(csub(t[1], t[3]));
    t[0] = cadd(s02, s13);
    t[2] = csub(s02, s13);
    t[1] = cadd(d02, d13);
    t[3] = csub(d02, d13);
  } else if constexpr (R == 3) {
    constexpr float s3 = (float)(0.86602540378443864676372317075294);
    float2 s = cadd(t[1], t[2]);
    float2 d = mul_i<SIGN>(csub(t[1], t[2]));
    float2 m = caxpy(t[0], -0.5f, s);
    t[0] = cadd(t[0], s);
    t[1] = caxpy(m, s3, d);
    t[2] = caxpy(m, -s3, d);
  } else if constexpr (R == 5) {
    constexpr float c1 = (float)ct::cos2pi(1, 5), c2 = (float)ct::cos2pi(2, 5);
    constexpr float s1 = (float)ct::sin2pi(1, 5), s2 = (float)ct::sin2pi(2, 5);
    float2 a14 = cadd(t[1], t[4]), b14 = mul_i<SIGN>(csub(t[1], t[4]));
    float2 a23 = cadd(t[2], t[3]), b23 = mul_i<SIGN>(csub(t[2], t[3]));
    float2 x0 = t[0];
    float2 m1 = caxpy(caxpy(x0, c1, a14), c2, a23);
    float2 m2 = caxpy(caxpy(x0, c2, a14), c1, a23);
    // w^q = cos + SIGN*i*sin: the odd part is sin * (SIGN*i)*(...), i.e. b14/b23
    float2 n1 = caxpy(make_float2(s1 * b14.x, s1 * b14.y), s2, b23);
    float2 n2 = caxpy(make_float2(s2 * b14.x, s2 * b14.y), -s1, b23);
    t[0] = cadd(x0, cadd(a14, a23));
    t[1] = cadd(m1, n1);
    t[4] = csub(m1, n1);
    t[2] = cadd(m2, n2);
    t[3] = csub(m2, n2);
  } else {
    float2 in[R];
    static_for<0, R>([&](auto r) { in[r] = t[r]; });
    static_for<0, R>([&](auto q) {
      float2 acc = in[0];
      static_for<1, R>([&](auto r) {
        constexpr int e = (decltype(r)::value * decltype(q)::value) % R;
        constexpr float wr = (float)ct::cos2pi(e, R);
        constexpr float wi = (float)(SIGN * ct::sin2pi(e, R));
        acc = cadd(acc, cmul(in[r], wr, wi));
      });
      t[q] = acc;
    });
  }
}

// out[k] = sum_{t<N} in[OFF + t*STRIDE] * exp(SIGN * 2 pi i t k / N), k < N.
template <int N, int STRIDE, int OFF, int SIGN, int LEN>
__device__ __forceinline__ void fft_dit(const float2 (&in)[LEN], float2 (&out)[N]) {
  if constexpr (N == 1) {
    out[0] = in[OFF];
  } else {
    constexpr int R = pick_radix(N);
    constexpr int M = N / R;
    if constexpr (M == 1) {
      float2 t[R];
      static_for<0, R>([&](auto r) { t[r] = in[OFF + decltype(r)::value * STRIDE]; });
      butterfly<R, SIGN>(t);
      static_for<0, R>([&](auto r) { out[r] = t[r]; });
    } else {
      float2 sub[R][M];
      static_for<0, R>([&](auto r) {
        fft_dit<M, STRIDE * R, OFF + decltype(r)::value * STRIDE, SIGN>(in, sub[decltype(r)::value]);
      });
      static_for<0, M>([&](auto k) {
        constexpr int kk = decltype(k)::value;
        float2 t[R];
        static_for<0, R>([&](auto r) {
          constexpr int rr = decltype(r)::value;
          constexpr int e = rr * kk;
          if constexpr (e % N == 0) {
            t[rr] = sub[rr][kk];
          } else {
            constexpr float wr = (float)ct::cos2pi(e, N);
            constexpr float wi = (float)(SIGN * ct::sin2pi(e, N));
            t[rr] = cmul(sub[rr][kk], wr, wi);
          }
        });
        butterfly<R, SIGN>(t);
        static_for<0, R>([&](auto q) { out[kk + decltype(q)::value * M] = t[q]; });
      });
    }
  }
}

template <int N, int SIGN>
__device__ __forceinline__ void fft(float2 (&x)[N]) {
  float2 y[N];
  fft_dit<N, 1, 0, SIGN>(x, y);
  static_for<0, N>([&](auto k) { x[k] = y[k]; });
}

}  // namespace nf3d
