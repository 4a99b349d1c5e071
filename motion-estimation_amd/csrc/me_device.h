// Device-side building blocks shared by the HIP kernels of this library:
// complex float2 arithmetic, compile-time twiddle tables, and small
// in-register DFTs composed from radix-2/3/4 butterflies.
//
// The FFTs here are textbook mixed-radix Cooley-Tukey (decimation in time),
// written for gfx950 wave64: each thread holds a short line segment in
// registers, and segments are exchanged through LDS between the two factor
// stages of a line.  Nothing here is derived from a vendor FFT library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace me {

// ---------------------------------------------------------------- complex
__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_scale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ float2 c_mulc(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -a.x * b.y));
}
// acc += a * b
__device__ __forceinline__ float2 c_fma(float2 a, float2 b, float2 acc) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(-a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(a.y, b.x, acc.y);
  return acc;
}
// acc += conj(a) * b
__device__ __forceinline__ float2 c_fmac(float2 a, float2 b, float2 acc) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(-a.y, b.x, acc.y);
  return acc;
}
__device__ __forceinline__ float2 c_conj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by i
__device__ __forceinline__ float2 c_muli(float2 a) { return make_float2(-a.y, a.x); }

// exp(i * 2*pi * f) for a reduced fraction f (|f| <= 1/2), FP32 result.
__device__ __forceinline__ float2 cis2pi(double f) {
  float s, c;
  sincospif((float)(2.0 * f), &s, &c);
  return make_float2(c, s);
}
// fractional part in [-1/2, 1/2]
__device__ __forceinline__ double frac_half(double v) { return v - rint(v); }

// ------------------------------------------------- compile-time trig tables
constexpr double kPi = 3.141592653589793238462643383279502884;

constexpr double ce_sin_poly(double x) {  // |x| <= pi/4, Taylor to ~1e-17
  double t = x, s = x;
  for (int n = 1; n < 12; ++n) {
    t *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
    s += t;
  }
  return s;
}
constexpr double ce_cos_poly(double x) {
  double t = 1.0, s = 1.0;
  for (int n = 1; n < 12; ++n) {
    t *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
    s += t;
  }
  return s;
}
// cos and sin of 2*pi*j/n, reduced by exact octant symmetry on the rational j/n.
struct CS {
  double c, s;
};
constexpr CS ce_cs_2pi(long j, long n) {
  j %= n;
  if (j < 0) j += n;
  // angle = 2*pi*j/n; quadrant q = floor(4j/n), remainder r in [0, n/4)
  long q = (4 * j) / n;
  long rem4 = 4 * j - q * n;  // 4*(j - q*n/4)
  double a = (kPi / 2.0) * (double)rem4 / (double)n;  // in [0, pi/2)
  double c = 0.0, s = 0.0;
  if (a <= kPi / 4) {
    c = ce_cos_poly(a);
    s = ce_sin_poly(a);
  } else {
    c = ce_sin_poly(kPi / 2 - a);
    s = ce_cos_poly(kPi / 2 - a);
  }
  switch (q & 3) {
    case 0: return CS{c, s};
    case 1: return CS{-s, c};
    case 2: return CS{-c, -s};
    default: return CS{s, -c};
  }
}

template <int N>
struct TwTab {
  float c[N];
  float s[N];
  constexpr TwTab() : c{}, s{} {
    for (int j = 0; j < N; ++j) {
      CS v = ce_cs_2pi(j, N);
      c[j] = (float)v.c;
      s[j] = (float)v.s;
    }
  }
};
template <int N>
__device__ constexpr TwTab<N> kTw{};

// v * exp(DIR * i * 2*pi * j / N) with j a compile-time constant after
// unrolling; trivial rotations are exact swaps.
template <int N, int DIR>
__device__ __forceinline__ float2 tw_const(float2 v, int j) {
  j %= N;
  if (j == 0) return v;
  if (2 * j == N) return make_float2(-v.x, -v.y);
  if (4 * j == N) return DIR < 0 ? make_float2(v.y, -v.x) : make_float2(-v.y, v.x);
  if (4 * j == 3 * N) return DIR < 0 ? make_float2(-v.y, v.x) : make_float2(v.y, -v.x);
  const float c = kTw<N>.c[j];
  const float s = DIR < 0 ? -kTw<N>.s[j] : kTw<N>.s[j];
  return c_mul(v, make_float2(c, s));
}

// ---------------------------------------------------- in-register DFTs
// RDFT<N, DIR>::run(v): unnormalised DFT, X[k] = sum_n v[n] exp(DIR*2*pi*i*n*k/N),
// natural order in and out.  DIR = -1 forward, +1 inverse.
template <int N, int DIR>
struct RDFT;

template <int DIR>
struct RDFT<1, DIR> {
  static __device__ __forceinline__ void run(float2*) {}
};
template <int DIR>
struct RDFT<2, DIR> {
  static __device__ __forceinline__ void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = c_add(a, b);
    v[1] = c_sub(a, b);
  }
};
template <int DIR>
struct RDFT<3, DIR> {
  static __device__ __forceinline__ void run(float2* v) {
    constexpr float h = 0.86602540378443864676f;  // sqrt(3)/2
    float2 a = v[0], b = v[1], c = v[2];
    float2 t1 = c_add(b, c);
    float2 t2 = make_float2(a.x - 0.5f * t1.x, a.y - 0.5f * t1.y);
    float2 d = c_sub(b, c);
    // DIR * i * h * d
    float2 rot = DIR < 0 ? make_float2(h * d.y, -h * d.x) : make_float2(-h * d.y, h * d.x);
    v[0] = c_add(a, t1);
    v[1] = c_add(t2, rot);
    v[2] = c_sub(t2, rot);
  }
};
template <int DIR>
struct RDFT<4, DIR> {
  static __device__ __forceinline__ void run(float2* v) {
    float2 s02 = c_add(v[0], v[2]), d02 = c_sub(v[0], v[2]);
    float2 s13 = c_add(v[1], v[3]), d13 = c_sub(v[1], v[3]);
    // DIR * i * d13
    float2 r = DIR < 0 ? make_float2(d13.y, -d13.x) : make_float2(-d13.y, d13.x);
    v[0] = c_add(s02, s13);
    v[2] = c_sub(s02, s13);
    v[1] = c_add(d02, r);
    v[3] = c_sub(d02, r);
  }
};

template <int N>
struct FirstFactor {
  static constexpr int value = (N % 4 == 0) ? 4 : ((N % 2 == 0) ? 2 : 3);
};

// Composite sizes: N = P*M, decimation in time.  P sub-DFTs of length M over
// the stride-P subsequences, twiddle exp(DIR 2 pi i p k / N), then M DFTs of
// length P.
template <int N, int DIR>
struct RDFT {
  static_assert(N == 6 || N == 8 || N == 12 || N == 16 || N == 24 || N == 32 || N == 48 || N == 64,
                "register DFT size");
  static __device__ __forceinline__ void run(float2* v) {
    constexpr int P = FirstFactor<N>::value;
    constexpr int M = N / P;
    float2 y[P][M];
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int m = 0; m < M; ++m) y[p][m] = v[P * m + p];
#pragma unroll
    for (int p = 0; p < P; ++p) RDFT<M, DIR>::run(y[p]);
#pragma unroll
    for (int k = 0; k < M; ++k) {
      float2 z[P];
#pragma unroll
      for (int p = 0; p < P; ++p) z[p] = tw_const<N, DIR>(y[p][k], p * k);
      RDFT<P, DIR>::run(z);
#pragma unroll
      for (int q = 0; q < P; ++q) v[k + M * q] = z[q];
    }
  }
};

// Factorisation of a line length N into (R1, R2): step 1 runs R2 threads of
// length-R1 register DFTs (each thread loads elements t, t+R2, ...), step 2
// runs R1 threads of length-R2 DFTs.  Threads per line T = max(R1, R2).
template <int N>
struct LineFactor;
#define ME_LINE_FACTOR(N_, R1_, R2_)          \
  template <>                                 \
  struct LineFactor<N_> {                     \
    static constexpr int R1 = R1_;            \
    static constexpr int R2 = R2_;            \
  };
ME_LINE_FACTOR(8, 2, 4)
ME_LINE_FACTOR(12, 3, 4)
ME_LINE_FACTOR(16, 4, 4)
ME_LINE_FACTOR(24, 3, 8)
ME_LINE_FACTOR(32, 4, 8)
ME_LINE_FACTOR(48, 6, 8)
ME_LINE_FACTOR(64, 8, 8)
ME_LINE_FACTOR(96, 12, 8)
ME_LINE_FACTOR(128, 8, 16)
ME_LINE_FACTOR(192, 12, 16)
ME_LINE_FACTOR(256, 16, 16)
ME_LINE_FACTOR(384, 24, 16)
ME_LINE_FACTOR(512, 32, 16)
#undef ME_LINE_FACTOR

// Calls F::template run<N>() for the supported line lengths; returns false if
// N is not supported.
template <class F>
inline bool dispatch_line_length(int n, F&& f) {
  switch (n) {
    case 8: f.template operator()<8>(); return true;
    case 12: f.template operator()<12>(); return true;
    case 16: f.template operator()<16>(); return true;
    case 24: f.template operator()<24>(); return true;
    case 32: f.template operator()<32>(); return true;
    case 48: f.template operator()<48>(); return true;
    case 64: f.template operator()<64>(); return true;
    case 96: f.template operator()<96>(); return true;
    case 128: f.template operator()<128>(); return true;
    case 192: f.template operator()<192>(); return true;
    case 256: f.template operator()<256>(); return true;
    case 384: f.template operator()<384>(); return true;
    case 512: f.template operator()<512>(); return true;
    default: return false;
  }
}

// Per-workgroup twiddle table exp(-2 pi i j / N), j in [0, N), filled once.
template <int N, int NT>
__device__ __forceinline__ void fill_twiddles(float2* tw) {
  // exp(-2 pi i j / N) from the compile-time table (rounded from FP64)
  for (int j = threadIdx.x; j < N; j += NT) tw[j] = make_float2(kTw<N>.c[j], -kTw<N>.s[j]);
}

}  // namespace me
