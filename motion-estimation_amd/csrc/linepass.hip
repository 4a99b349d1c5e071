// Line-pass kernel: for every line of a volume along one axis,
//   forward FFT -> per-(output,input) diagonal phase -> inverse FFT,
// entirely in registers + LDS, one HBM read of each input and one write of
// each output.  This is the building block of the FFT-shear rigid transform
// (transform.py:243-504): every shear step "fft along an axis, multiply by
// exp(i c r_b k_a), ifft" (transform.py:279-301) is one pass, translation
// factors (transform.py:308-312) fold into passes along the same axis, and
// the rotation-derivative product rule (transform.py:433-474) becomes passes
// with two inputs and/or two outputs.
//
// Thread layout: a line of N = R1*R2 points is owned by T = max(R1, R2)
// threads; step 1 runs R2 threads each holding the R1 points t, t+R2, ...
// (a register DFT of length R1), one LDS exchange, step 2 runs R1 threads of
// length-R2 register DFTs.  The inverse retraces the same path, so each thread
// writes exactly the addresses it read (passes may run in place).
// A workgroup holds L lines.  For a strided line axis the lines are
// consecutive along the unit-stride axis (lanes line-fastest, L*8-byte
// segments); for a unit-stride line axis, lanes run along the line.
#include "me_device.h"
#include "me_internal.h"

#include <map>
#include <mutex>
#include <tuple>

namespace me {

// Lines per workgroup LL: 16 by default for both layouts (strided lanes
// line-fastest: 128-byte contiguous segments per line position).  The plain
// single-buffer pass (one input, one output, FFT-phase-inverse FFT) is also
// built for 8 and 32 lines, picked per geometry by the tuning table below
// (me_tune_line_passes times the variants on the caller's buffers).
constexpr int kLinesDefault = 16;

// A pass reads every line value once and writes it once: streaming
// (nontemporal) loads and stores.  Together they took the 10-pass transform
// of 64 volumes at 256x256x192 from 24.0 to 23.2 ms; either one alone changed
// nothing.
typedef float f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 stream_load(const float2* p) {
  const f2v w = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
  return make_float2(w.x, w.y);
}
__device__ __forceinline__ void stream_store(float2* p, float2 v) {
  f2v w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<f2v*>(p));
}

template <int N, bool CONTIG, int LL = kLinesDefault>
struct LineT {
  static constexpr int R1 = LineFactor<N>::R1;
  static constexpr int R2 = LineFactor<N>::R2;
  static constexpr int T = R1 > R2 ? R1 : R2;
  static constexpr int L = LL;
  static constexpr int NT = T * L;
};

// acc[k2] += scale * phi(k) * X[k2] for one (output, input) pair on the k
// values owned by a step-2 thread, k = t + R1*k2.  The phase exp(i k_a beta)
// is factored as exp(2 pi i beta t/N) * exp(2 pi i beta R1/N)^k2, times
// exp(-2 pi i beta) for k >= N/2 (k_a = 2 pi (k - N)/N there), with each
// factor reduced in FP64 before the FP32 sincos.  For even R2 (N/2 = R1*R2/2
// and t < R1) k >= N/2 exactly when k2 >= R2/2, so the wrap is applied once
// at that compile-time position.
template <int N, int R1, int R2, bool DERIV>
__device__ __forceinline__ void phase_loop(float2 cur, float2 step, float2 wrap, float gam, int t,
                                           const float2 (&X)[R2], float2 (&acc)[R2]) {
  constexpr bool kEven = (R2 % 2) == 0;
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) {
    const int k = t + R1 * k2;
    bool w;
    float2 e;
    if (kEven) {
      w = k2 >= R2 / 2;
      if (k2 == R2 / 2) cur = c_mul(cur, wrap);
      e = cur;
    } else {
      w = 2 * k >= N;
      e = w ? c_mul(cur, wrap) : cur;
    }
    if (DERIV) {
      const float kap = (float)(k - (w ? N : 0)) / (float)N;
      e = c_muli(c_scale(e, gam * kap));
    }
    acc[k2] = c_fma(X[k2], e, acc[k2]);
    if (k2 + 1 < R2) cur = c_mul(cur, step);
  }
}

template <int N, int R1, int R2>
__device__ __forceinline__ void phase_acc(const PhaseSpec& P, const double* cf, double rb, int t, float scale,
                                          const float2 (&X)[R2], float2 (&acc)[R2]) {
  const double c = P.ic >= 0 ? (double)P.sc * cf[P.ic] : 0.0;
  const double d = P.id >= 0 ? (double)P.sd * cf[P.id] : 0.0;
  const double beta = c * rb + d;  // angle = k_a * beta
  // compile-time reciprocals instead of FP64 divisions (R1/N = 1/R2)
  constexpr double kInvN = 1.0 / (double)N;
  constexpr double kInvR2 = 1.0 / (double)R2;
  const float2 cur = c_scale(cis2pi(frac_half(beta * ((double)t * kInvN))), scale);
  const float2 step = cis2pi(frac_half(beta * kInvR2));
  const float2 wrap = cis2pi(frac_half(-beta));
  if (P.deriv) {  // uniform: one of two unrolled loops
    const double g = P.ig >= 0 ? (double)P.sg * cf[P.ig] : 0.0;
    const float gam = (float)(2.0 * kPi * (g * rb + (double)P.h));
    phase_loop<N, R1, R2, true>(cur, step, wrap, gam, t, X, acc);
  } else {
    phase_loop<N, R1, R2, false>(cur, step, wrap, 0.f, t, X, acc);
  }
}

// MODE 0: FFT, phase, inverse FFT; 1: forward FFT only (natural-order
// spectrum out); 2: inverse FFT only (natural-order spectrum in).
// REDUCE: sum the result over the batch and write batch entry 0 only.
// DOT: the dot epilogue (LineArgs::dot_w) instead of storing the output.
template <int N, bool CONTIG, int NIN, int NOUT, int MODE, bool REDUCE, int LL = kLinesDefault, bool DOT = false>
__global__ void __launch_bounds__((LineT<N, CONTIG, LL>::NT)) k_linepass(LineArgs A) {
  static_assert(!DOT || (NOUT == 1 && MODE == 0 && !REDUCE), "dot epilogue: single-output MODE 0 passes");
  using LT = LineT<N, CONTIG, LL>;
  constexpr int R1 = LT::R1;
  constexpr int R2 = LT::R2;
  constexpr int T = LT::T;
  constexpr int L = LT::L;
  constexpr int NT = LT::NT;
  __shared__ float2 lds[L * R1 * (R2 + 1)];
  __shared__ float2 tw[N];
  fill_twiddles<N, NT>(tw);

  const int tid = threadIdx.x;
  const int t = CONTIG ? (tid % T) : (tid / L);
  const int l = CONTIG ? (tid / T) : (tid % L);
  int u, v;
  bool valid;
  if (CONTIG) {
    const long long g = (long long)blockIdx.x * L + l;
    valid = g < (long long)A.n_u * A.n_v;
    u = valid ? (int)(g % A.n_u) : 0;
    v = valid ? (int)(g / A.n_u) : 0;
  } else {
    const int tiles_u = (A.n_u + L - 1) / L;
    v = blockIdx.x / tiles_u;
    u = (blockIdx.x % tiles_u) * L + l;
    valid = u < A.n_u;
    if (!valid) u = 0;
  }
  const long long line_off = (long long)u * A.st_u + (long long)v * A.st_v;
  auto lidx = [&](int k1, int n2) -> int {
    return CONTIG ? (l * (R1 * (R2 + 1)) + k1 * (R2 + 1) + n2) : (k1 * ((R2 + 1) * L) + n2 * L + l);
  };
  const double ru = A.r_u ? A.r_u[u] : 0.0;
  const double rv = A.r_v ? A.r_v[v] : 0.0;
  // padding lines (valid == false) read line 0, so every load is
  // unconditional and a thread's loads are all in flight together; lines are
  // independent and a padding line's results are never stored
  // MODE 0 folds the 1/N of the FFT pair into the phase accumulation;
  // MODE 1/2 scale the stored values by N^-1/2
  const float ph_scale = MODE == 0 ? 1.0f / (float)N : 1.0f;
  const float out_scale = MODE == 0 ? 1.0f : rsqrtf((float)N);
  __syncthreads();

  double dotv = 0.0;  // dot epilogue: this thread's share of Re(conj(w) out)
  const int b_begin = REDUCE ? 0 : (int)blockIdx.y;
  const int b_end = REDUCE ? A.nbatch : (int)blockIdx.y + 1;
  float2 red[REDUCE ? R1 : 1];
#pragma unroll
  for (int i = 0; i < (REDUCE ? R1 : 1); ++i) red[i] = make_float2(0.f, 0.f);

  // The line's values of every input (MODE 0/1) are loaded before any of
  // them is transformed, so the second input's loads are not held back by the
  // barriers of the first one's transform.  (Every value is read before any
  // is written: a workgroup owns its lines in every buffer.)  A REDUCE pass
  // loads each batch entry at the top of its iteration: loading the next one
  // during the current one's transform costs it occupancy (157 -> 222 VGPRs)
  // and made it slower (1.53 -> 1.65 ms at 64 volumes of 256x256x192).
  constexpr int NRAW = MODE == 2 ? 1 : R1;
  float2 raw[NIN][NRAW];
  auto load_raw = [&](int bb) {
    if (MODE != 2 && t < R2) {
      const long long step = (long long)R2 * A.st_a;
#pragma unroll
      for (int i = 0; i < NIN; ++i) {
        const float2* st = A.in[i] + (long long)bb * A.in_bs[i] + line_off + (long long)t * A.st_a;
#pragma unroll
        for (int n1 = 0; n1 < NRAW; ++n1) raw[i][n1] = stream_load(st + n1 * step);
      }
    }
  };

  for (int b = b_begin; b < b_end; ++b) {
    const double* cf = A.coef ? A.coef + (long long)b * A.ncoef : nullptr;
    load_raw(b);
    // ---------------- forward transforms of every input
    float2 X[NIN][R2];
#pragma unroll
    for (int i = 0; i < NIN; ++i) {
      const float2* src = A.in[i] + (long long)b * A.in_bs[i] + line_off;
      if (MODE == 2) {
        if (t < R1) {
          const float2* st = src + (long long)t * A.st_a;
          const long long step = (long long)R1 * A.st_a;
#pragma unroll
          for (int k2 = 0; k2 < R2; ++k2) X[i][k2] = st[k2 * step];
        }
        continue;
      }
      if (t < R2) {
        float2 a[R1];
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) a[n1] = raw[i][n1 < NRAW ? n1 : 0];
        RDFT<R1, -1>::run(a);
#pragma unroll
        for (int k1 = 1; k1 < R1; ++k1) a[k1] = c_mul(a[k1], tw[(t * k1) % N]);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) lds[lidx(k1, t)] = a[k1];
      }
      __syncthreads();
      if (t < R1) {
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) X[i][n2] = lds[lidx(t, n2)];
        RDFT<R2, -1>::run(X[i]);
      }
      __syncthreads();
    }

    // ---------------- per output: combine inputs with their phases, inverse
#pragma unroll
    for (int o = 0; o < NOUT; ++o) {
      float2 acc[R2];
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) acc[k2] = make_float2(0.f, 0.f);
      if (t < R1) {
#pragma unroll
        for (int i = 0; i < NIN; ++i) {
          const PhaseSpec& P = A.ph[o][i];
          if (!P.active) continue;
          if (P.ic < 0 && P.id < 0 && !P.deriv) {  // unit phase
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) {
              acc[k2].x = fmaf(ph_scale, X[i][k2].x, acc[k2].x);
              acc[k2].y = fmaf(ph_scale, X[i][k2].y, acc[k2].y);
            }
          } else {
            const double rb = P.slot == 0 ? ru : (P.slot == 1 ? rv : 0.0);
            phase_acc<N, R1, R2>(P, cf, rb, t, ph_scale, X[i], acc);
          }
        }
      }
      float2* dst = A.out[o] + (long long)(REDUCE ? 0 : b) * A.out_bs[o] + line_off;
      if (MODE == 1) {  // forward only: store natural-order spectrum
        if (t < R1 && valid) {
          float2* dt = dst + (long long)t * A.st_a;
          const long long step = (long long)R1 * A.st_a;
#pragma unroll
          for (int k2 = 0; k2 < R2; ++k2) {
            float2 val = c_scale(acc[k2], out_scale);
            if (A.accumulate) val = c_add(val, dt[k2 * step]);
            dt[k2 * step] = val;
          }
        }
        continue;
      }
      if (t < R1) {
        RDFT<R2, +1>::run(acc);
#pragma unroll
        for (int n2 = 1; n2 < R2; ++n2) acc[n2] = c_mulc(acc[n2], tw[(n2 * t) % N]);
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) lds[lidx(t, n2)] = acc[n2];
      }
      __syncthreads();
      if (t < R2) {
        float2 a[R1];
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) a[k1] = lds[lidx(k1, t)];
        RDFT<R1, +1>::run(a);
        if (REDUCE) {
#pragma unroll
          for (int n1 = 0; n1 < R1; ++n1) red[n1 < (REDUCE ? R1 : 1) ? n1 : 0] = c_add(red[n1 < (REDUCE ? R1 : 1) ? n1 : 0], a[n1]);
        } else if (DOT) {
          if (valid) {
            const float2* wt = A.dot_w + (long long)b * A.dot_w_bs + line_off + (long long)t * A.st_a;
            const long long step = (long long)R2 * A.st_a;
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) {
              const float2 w = wt[n1 * step];
              dotv += (double)w.x * a[n1].x + (double)w.y * a[n1].y;
            }
          }
        } else if (valid) {
          float2* dt = dst + (long long)t * A.st_a;
          const long long step = (long long)R2 * A.st_a;
          if (!A.mask && !A.accumulate) {  // the common case: stores only, no per-element tests
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) stream_store(dt + n1 * step, MODE == 0 ? a[n1] : c_scale(a[n1], out_scale));
          } else {
            const unsigned char* mk =
                A.mask ? A.mask + (long long)b * A.mask_bs + line_off + (long long)t * A.st_a : nullptr;
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) {
              float2 val = MODE == 0 ? a[n1] : c_scale(a[n1], out_scale);
              if (mk && !mk[n1 * step]) val = make_float2(0.f, 0.f);
              if (A.accumulate) val = c_add(val, dt[n1 * step]);
              dt[n1 * step] = val;
            }
          }
        }
      }
      __syncthreads();
    }
  }
  if (DOT) {  // one partial per workgroup
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dotv += __shfl_down(dotv, o, 64);
    double* sh = reinterpret_cast<double*>(lds);  // free after the last barrier
    if ((tid & 63) == 0) sh[tid >> 6] = dotv;
    __syncthreads();
    if (tid == 0) {
      double sum = 0.0;
      for (int w = 0; w < (NT + 63) / 64; ++w) sum += sh[w];
      A.dot_partial[(long long)blockIdx.y * gridDim.x + blockIdx.x] = sum;
    }
  }
  if (REDUCE && t < R2 && valid) {
    float2* dt = A.out[0] + line_off + (long long)t * A.st_a;
    const long long step = (long long)R2 * A.st_a;
    const unsigned char* mk = A.mask ? A.mask + line_off + (long long)t * A.st_a : nullptr;
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
      float2 val = red[n1 < (REDUCE ? R1 : 1) ? n1 : 0];  // MODE 0: scaled in the phase step
      if (mk && !mk[n1 * step]) val = make_float2(0.f, 0.f);
      if (A.accumulate) val = c_add(val, dt[n1 * step]);
      dt[n1 * step] = val;
    }
  }
}

namespace {
// Tuning table: lines per workgroup of the plain pass, keyed by the line
// geometry of a launch.  Process-wide, filled by me_tune_line_passes or
// me_line_tuning_set; absent keys use kLinesDefault.
struct LineKey {
  int n, contig, n_u, n_v;
  long long st_a, st_u, st_v;
  bool operator<(const LineKey& o) const {
    return std::tie(n, contig, n_u, n_v, st_a, st_u, st_v) <
           std::tie(o.n, o.contig, o.n_u, o.n_v, o.st_a, o.st_u, o.st_v);
  }
};
std::mutex g_tune_mu;
std::map<LineKey, int> g_tune;

int tuned_lines(const LineKey& k) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(k);
  return it == g_tune.end() ? kLinesDefault : it->second;
}

template <int N, bool CONTIG, int LL>
constexpr bool lines_fit() {
  using LT = LineT<N, CONTIG, LL>;
  // at most 512 threads: 1024-thread variants of the long lines spill
  return LT::NT <= 512 && (LT::L * LT::R1 * (LT::R2 + 1) + N) * 8 <= 160 * 1024;
}

template <bool CONTIG, int NIN, int NOUT, int MODE, bool REDUCE>
struct LineLaunch {
  const LineArgs* args;
  dim3 grid;
  hipStream_t stream;
  int lines;
  long long* grid_x;
  template <int N, int LL, bool DOT = false>
  void go() {
    using LT = LineT<N, CONTIG, LL>;
    const long long nblk = CONTIG ? ((long long)args->n_u * args->n_v + LT::L - 1) / LT::L
                                  : (long long)((args->n_u + LT::L - 1) / LT::L) * args->n_v;
    dim3 g = grid;
    g.x = (unsigned)nblk;
    if (grid_x) *grid_x = nblk;
    hipLaunchKernelGGL((k_linepass<N, CONTIG, NIN, NOUT, MODE, REDUCE, LL, DOT>), g, dim3(LT::NT), 0, stream,
                       *args);
  }
  template <int N>
  void operator()() {
    if constexpr (NIN == 1 && NOUT == 1 && MODE == 0 && !REDUCE) {
      if (args->dot_w) return go<N, kLinesDefault, true>();
      if constexpr (lines_fit<N, CONTIG, 8>())
        if (lines == 8) return go<N, 8>();
      if constexpr (lines_fit<N, CONTIG, 32>())
        if (lines == 32) return go<N, 32>();
    }
    go<N, kLinesDefault>();
  }
};

template <bool CONTIG, int NIN, int NOUT, int MODE, bool REDUCE>
bool launch_n(int n, const LineArgs& a, dim3 grid, hipStream_t s, int lines, long long* grid_x) {
  LineLaunch<CONTIG, NIN, NOUT, MODE, REDUCE> f{&a, grid, s, lines, grid_x};
  return dispatch_line_length(n, f);
}

template <bool CONTIG>
bool launch_io(int n, int nin, int nout, const LineArgs& a, dim3 grid, hipStream_t s, int lines,
               long long* gx) {
  if (a.mode == 1) return nin == 1 && nout == 1 && launch_n<CONTIG, 1, 1, 1, false>(n, a, grid, s, lines, gx);
  if (a.mode == 2) return nin == 1 && nout == 1 && launch_n<CONTIG, 1, 1, 2, false>(n, a, grid, s, lines, gx);
  if (a.reduce) return nin == 1 && nout == 1 && launch_n<CONTIG, 1, 1, 0, true>(n, a, grid, s, lines, gx);
  if (nin == 1 && nout == 1) return launch_n<CONTIG, 1, 1, 0, false>(n, a, grid, s, lines, gx);
  if (nin == 1 && nout == 2) return launch_n<CONTIG, 1, 2, 0, false>(n, a, grid, s, lines, gx);
  if (nin == 2 && nout == 2) return launch_n<CONTIG, 2, 2, 0, false>(n, a, grid, s, lines, gx);
  if (nin == 2 && nout == 1) return launch_n<CONTIG, 2, 1, 0, false>(n, a, grid, s, lines, gx);
  return false;
}
}  // namespace

// Tile axes of a pass along ``axis``: au runs across the lines of a
// workgroup, av across workgroups.  contig: the line axis has unit stride.
static void line_layout(const VolGeom& g, int axis, bool& contig, int& au, int& av) {
  const int o1 = axis == 0 ? 1 : 0;
  const int o2 = axis == 2 ? 1 : 2;
  if (g.st[axis] == 1) {
    contig = true;
    au = o1;
    av = o2;
    if (g.st[o2] < g.st[o1]) {
      au = o2;
      av = o1;
    }
  } else {
    contig = false;
    if (g.st[o1] == 1) {
      au = o1;
      av = o2;
    } else if (g.st[o2] == 1) {
      au = o2;
      av = o1;
    } else {
      au = g.st[o1] <= g.st[o2] ? o1 : o2;
      av = au == o1 ? o2 : o1;
    }
  }
}

static LineKey line_key(const VolGeom& g, int axis) {
  bool contig;
  int au, av;
  line_layout(g, axis, contig, au, av);
  return LineKey{g.n[axis], contig ? 1 : 0, g.n[au], g.n[av], g.st[axis], g.st[au], g.st[av]};
}

// Whether the plain pass of this line length and layout is built for `lines`.
struct LinesBuilt {
  bool contig;
  int lines;
  bool ok = false;
  template <int N>
  void operator()() {
    if (lines == kLinesDefault) ok = true;
    if (lines == 8) ok = contig ? lines_fit<N, true, 8>() : lines_fit<N, false, 8>();
    if (lines == 32) ok = contig ? lines_fit<N, true, 32>() : lines_fit<N, false, 32>();
  }
};
static bool lines_built(int n, bool contig, int lines) {
  LinesBuilt f{contig, lines};
  return dispatch_line_length(n, f) && f.ok;
}

int launch_linepass(const VolGeom& g, int axis, int nin, int nout, LineArgs A, const HostPhase ph[2][2],
                    hipStream_t stream, long long* grid_x) {
  ME_REQUIRE(axis >= 0 && axis < 3, "line pass: axis out of range");
  ME_REQUIRE(nin >= 1 && nin <= 2 && nout >= 1 && nout <= 2, "line pass: 1 or 2 inputs/outputs");
  ME_REQUIRE(!(A.reduce && (nout != 1 || nin != 1)), "line pass: batch reduction needs one input and output");
  ME_REQUIRE(A.mode == 0 || (nin == 1 && nout == 1 && !A.reduce), "line pass: FFT-only modes are single-buffer");
  ME_REQUIRE(!A.dot_w || (A.mode == 0 && nin == 1 && nout == 1 && !A.reduce && A.dot_partial),
             "line pass: the dot epilogue needs a one-input, one-output MODE 0 pass and a partials buffer");
  ME_REQUIRE(A.nbatch >= 1, "line pass: empty batch");
  int au, av;
  bool contig;
  line_layout(g, axis, contig, au, av);
  A.st_a = g.st[axis];
  A.st_u = g.st[au];
  A.st_v = g.st[av];
  A.n_u = g.n[au];
  A.n_v = g.n[av];
  A.r_u = g.r[au];
  A.r_v = g.r[av];
  for (int o = 0; o < 2; ++o)
    for (int i = 0; i < 2; ++i) {
      const HostPhase& h = ph[o][i];
      PhaseSpec& p = A.ph[o][i];
      p.active = (o < nout && i < nin && h.active) ? 1 : 0;
      p.ic = h.ic;
      p.sc = h.sc;
      p.id = h.id;
      p.sd = h.sd;
      p.ig = h.ig;
      p.sg = h.sg;
      p.h = h.h;
      p.deriv = h.deriv;
      p.slot = -1;
      if (p.active && h.b_axis >= 0) {
        ME_REQUIRE(h.b_axis != axis, "line pass: phase cannot depend on the line axis");
        p.slot = h.b_axis == au ? 0 : 1;
        ME_REQUIRE(g.r[h.b_axis] != nullptr, "line pass: missing grid vector");
      }
      if (p.active && (p.ic >= 0 || p.id >= 0 || p.ig >= 0))
        ME_REQUIRE(A.coef != nullptr, "line pass: missing coefficient table");
    }
  // grid.x (line tiles) is set per line length by the launcher
  const long long nblk = contig ? ((long long)A.n_u * A.n_v + 7) / 8 : (long long)A.n_u * A.n_v;
  ME_REQUIRE(nblk > 0 && nblk < (1LL << 31), "line pass: grid too large");
  dim3 grid(1u, A.reduce ? 1u : (unsigned)A.nbatch);
  ME_REQUIRE(A.reduce || A.nbatch <= 65535, "line pass: batch too large");
  const int lines = tuned_lines(LineKey{g.n[axis], contig ? 1 : 0, A.n_u, A.n_v, A.st_a, A.st_u, A.st_v});
  const bool ok = contig ? launch_io<true>(g.n[axis], nin, nout, A, grid, stream, lines, grid_x)
                         : launch_io<false>(g.n[axis], nin, nout, A, grid, stream, lines, grid_x);
  ME_REQUIRE(ok, "line pass: unsupported line length " + std::to_string(g.n[axis]) +
                     " (supported: 8,12,16,24,32,48,64,96,128,192,256,384,512)");
  ME_TRY_HIP(hipGetLastError());
  return 0;
}


long long linepass_max_tiles(const VolGeom& g, int axis) {
  bool contig;
  int au, av;
  line_layout(g, axis, contig, au, av);
  const long long lines = 8;  // the fewest lines per workgroup of any variant
  return contig ? ((long long)g.n[au] * g.n[av] + lines - 1) / lines
                : ((long long)(g.n[au] + lines - 1) / lines) * g.n[av];
}

static int tune_set(const VolGeom& g, int axis, int lines) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  const LineKey k = line_key(g, axis);
  if (lines == 0)
    g_tune.erase(k);
  else
    g_tune[k] = lines;
  return 0;
}

}  // namespace me

using namespace me;

static bool geom_from(const int* n, const long long* st, VolGeom& g) {
  if (!n || !st) return false;
  for (int a = 0; a < 3; ++a) {
    if (n[a] < 1 || st[a] < 1) return false;
    g.n[a] = n[a];
    g.st[a] = st[a];
    g.r[a] = nullptr;
  }
  return true;
}

extern "C" {

int me_line_tuning_set(const int* n, const long long* st, int axis, int lines) {
  VolGeom g;
  ME_REQUIRE(geom_from(n, st, g) && axis >= 0 && axis < 3, "line_tuning_set: bad geometry");
  bool contig;
  int au, av;
  line_layout(g, axis, contig, au, av);
  ME_REQUIRE(lines == 0 || lines_built(n[axis], contig, lines),
             "line_tuning_set: lines per workgroup not built for this line length (8, 16 or 32)");
  return tune_set(g, axis, lines);
}

int me_line_tuning_get(const int* n, const long long* st, int axis) {
  VolGeom g;
  if (!geom_from(n, st, g) || axis < 0 || axis > 2) return 0;
  return tuned_lines(line_key(g, axis));
}

void me_line_tuning_clear(void) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
}

int me_tune_line_passes(const void* x, void* y, long long bstride, int nbatch, const int* n, const long long* st,
                        int reps, int* chosen, float* times_ms, hipStream_t stream) {
  VolGeom g;
  ME_REQUIRE(x && y && nbatch >= 1 && reps >= 1 && geom_from(n, st, g), "tune_line_passes: bad argument");
  static const int kCand[3] = {8, kLinesDefault, 32};
  hipEvent_t e0, e1;
  ME_TRY_HIP(hipEventCreate(&e0));
  ME_TRY_HIP(hipEventCreate(&e1));
  int rc = 0;
  for (int axis = 0; axis < 3 && rc == 0; ++axis) {
    bool contig;
    int au, av;
    line_layout(g, axis, contig, au, av);
    LineArgs A{};
    A.nbatch = nbatch;
    A.mode = 0;
    A.in[0] = (const float2*)x;
    A.in_bs[0] = bstride;
    A.out[0] = (float2*)y;
    A.out_bs[0] = bstride;
    HostPhase ph[2][2];
    ph[0][0].active = true;  // unit phase: FFT, scale, inverse FFT
    int best = kLinesDefault;
    float best_t = 0.f;
    for (int c = 0; c < 3 && rc == 0; ++c) {
      const int L = kCand[c];
      if (times_ms) times_ms[axis * 3 + c] = -1.f;
      if (!lines_built(n[axis], contig, L)) continue;
      tune_set(g, axis, L);
      rc = launch_linepass(g, axis, 1, 1, A, ph, stream);  // warm-up
      if (rc) break;
      if (hipEventRecord(e0, stream) != hipSuccess) {
        set_error("tune_line_passes: hipEventRecord failed");
        rc = 1;
      }
      for (int r = 0; r < reps && rc == 0; ++r) rc = launch_linepass(g, axis, 1, 1, A, ph, stream);
      float ms = 0.f;
      if (rc == 0 && (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                      hipEventElapsedTime(&ms, e0, e1) != hipSuccess)) {
        set_error("tune_line_passes: timing events failed");
        rc = 1;
      }
      if (rc) break;
      ms /= (float)reps;
      if (times_ms) times_ms[axis * 3 + c] = ms;
      if (best_t == 0.f || ms < best_t) {
        best_t = ms;
        best = L;
      }
    }
    if (rc == 0) {
      tune_set(g, axis, best);
      if (chosen) chosen[axis] = best;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

}  // extern "C"
