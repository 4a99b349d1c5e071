// Per-shot SENSE encoding E_s = M_s F C and its adjoint, in the hybrid
// (k0, k1, x2) space.
//
// The reference computes, per shot s and coil c (image_estimation.py:94-128,
// transform_estimation.py:42-45, joint_recon.py:112-113)
//     E_s u = M_s * fftn(S_c * u)   over all three axes, with a dense boolean
// mask M_s(k0, k1) broadcast along the readout axis 2.  Because the mask does
// not depend on k2 and the 3-D FFT is orthonormal and separable, the readout
// FFT can be moved onto the measured data once (hybrid space
// y_h = ifft_2(y), a norm- and inner-product-preserving change of basis), so
// only the 2-D transform over (x0, x1) and only the sampled (k0, k1) lines are
// ever computed:
//     Y_s[x2][c][w] = (N0 N1)^-1/2 sum_{x0,x1} e^{-2 pi i (k0 x0/N0 + k1 x1/N1)} S_c u
// for the samples w = (k0, k1) of shot s.  Forward: a workgroup owns one
// readout plane x2 and 16 coils, streams the N0 rows of S_c * u through a
// row FFT along axis 1 (registers + LDS) and accumulates the sparse DFT along
// axis 0 for its samples in registers, so each map and image element is read
// from HBM once and only the samples are written.  Adjoint: a workgroup owns
// a block of rows of one plane, stages a coil group's samples in LDS, builds
// each row's spectrum by a column-sorted scatter, inverse row FFT, multiplies
// by conj(S_c), and sums the coils in registers.
//
// Volumes (image u, maps S_c) are in plane-major layout: element (x0, x1, x2)
// at (x2*N0 + x0)*N1 + x1.  Samples of a shot are sorted by (k1, k0).
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "me_device.h"
#include "me_internal.h"
#include "../../include/motion_estimation_amd.h"

namespace me {

constexpr int kMaxCol = 512;  // maximum N0 (column length, LDS twiddle table)
constexpr int kAdjChunk = 512;  // samples per coil staged in LDS by the adjoint

template <int NR>
struct RowT {
  static constexpr int R1 = LineFactor<NR>::R1;
  static constexpr int R2 = LineFactor<NR>::R2;
  static constexpr int T = R1 > R2 ? R1 : R2;  // threads per row
};

// Forward-encoding variants: G coils per workgroup (one row of T threads per
// coil) and SPT samples accumulated per thread.  A shot whose samples exceed
// SPT * T * G is split into chunks that each stream all rows again.  Every
// variant computes each sample with the same operations (bitwise the same
// samples); only the grouping of the FP64 norm partials differs.
struct FwdVariant {
  int G, SPT;
};
constexpr FwdVariant kFwdVariants[] = {{16, 2}, {8, 2}, {8, 4}, {4, 4}};
constexpr int kNumFwdVariants = 4;

template <int NR, int G>
struct FwdT {
  static constexpr int R1 = RowT<NR>::R1, R2 = RowT<NR>::R2, T = RowT<NR>::T;
  static constexpr int NT = T * G;
};

struct EncArgs {
  const float2* u;   // image batch (plane-major)
  long long u_bs;    // batch stride (0: same image for every shot)
  const float2* maps;
  int ncoils;
  int n0, n1, n2;
  const int* sk0;   // sample k0 (global list)
  const int* sk1;   // sample k1
  const int* off;   // per batch entry: first sample in the global list
  const int* cnt;   // per batch entry: number of samples
  const int* colptr;  // per batch entry: (n1+1) column pointers, relative to off
  float2* y;          // gathered output (may be null when only the norm is wanted)
  const float2* meas;  // gathered measured data to subtract (null: none)
  double* partial;     // per-workgroup FP64 partial sums of |y - meas|^2 (null: none)
  int ncg;             // coil groups
  int nchunk;          // sample chunks per shot
  int blocks_per_shot;
  float2* out;         // adjoint output batch (plane-major)
  long long out_bs;
  int rows_per_wg;
  int max_cnt;
  int nbatch;
  int store_raw;  // forward: store Y itself even when meas is given (norms are still of Y - meas)
};

// Workgroups are dealt to the 8 XCDs round robin (workgroup i to XCD i % 8).
// This maps them so that XCD x runs one contiguous range of the logical
// order, so the shots of one readout plane (consecutive in the logical
// order) share that XCD's L2 for the plane's map rows.
__device__ __forceinline__ int xcd_block(int i, int n) {
  const int per = n / 8, rem = n % 8;
  const int x = i % 8, k = i / 8;
  return x * per + min(x, rem) + k;
}

__device__ __forceinline__ double block_sum_f64(double v, double* sh) {
  // wave64 shuffle reduction, then one value per wave through LDS
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) s += sh[i];
  }
  return s;
}

// FULL: the launch's largest shot fills whole chunks (max_cnt a multiple of
// SPT * NT), so the sample slots are accumulated without a per-lane test (a
// slot past its shot's count has k0 = k1 = 0 and an accumulator that is never
// stored) and share one block, their spectrum reads in flight together;
// otherwise only the lanes whose slot holds a sample accumulate.  Both give
// bitwise the same samples.
template <int NR, int SPT, int G, bool FULL>
__global__ void __launch_bounds__((FwdT<NR, G>::NT)) k_encode_fwd(EncArgs A) {
  constexpr int R1 = FwdT<NR, G>::R1, R2 = FwdT<NR, G>::R2, T = FwdT<NR, G>::T, NT = FwdT<NR, G>::NT;
  // exchange rows padded to XS (even, so each row starts 16-byte aligned and
  // the step-2 reads go in pairs)
  constexpr int XS = R2 + 2;
  static_assert(R2 % 2 == 0, "paired exchange reads need an even R2");
  __shared__ __align__(16) float2 xch[G * R1 * XS];
  // row spectrum, coil-fastest per column k1 (stride ZS = G + 2 keeps 16-byte
  // alignment for paired coil reads and spreads the step-2 writes over banks)
  constexpr int ZS = G + 2;
  __shared__ __align__(16) float2 zrow[NR * ZS];
  __shared__ float2 twr[NR];
  __shared__ float2 twc[kMaxCol];
  __shared__ double red[16];

  const int n0 = A.n0, n1 = A.n1, n2 = A.n2;
  // block index: shot fastest, so the workgroups of all shots for one
  // (plane, coil group) run together and share the map rows through L2
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  int bx = lb;
  const int b = bx % A.nbatch;
  bx /= A.nbatch;
  const int x2 = bx % n2;
  bx /= n2;
  const int cg = bx % A.ncg;
  const int chunk = bx / A.ncg;
  const int tid = threadIdx.x;
  const int t = tid % T, g = tid / T;
  const int c = cg * G + g;
  const bool cvalid = c < A.ncoils;
  const int cnt = A.cnt[b], off = A.off[b];

  fill_twiddles<NR, NT>(twr);
  for (int j = tid; j < n0; j += NT) {
    float s, co;
    const int jj = (2 * j > n0) ? j - n0 : j;
    sincospif(-2.0f * (float)jj / (float)n0, &s, &co);
    twc[j] = make_float2(co, s);
  }
  int mk0[SPT], mk1[SPT];
  bool mv[SPT];
#pragma unroll
  for (int j = 0; j < SPT; ++j) {
    const int w = chunk * (SPT * NT) + j * NT + tid;
    mv[j] = w < cnt;
    mk0[j] = mv[j] ? A.sk0[off + w] : 0;
    mk1[j] = mv[j] ? A.sk1[off + w] : 0;
  }
  float2 acc[SPT][G];
#pragma unroll
  for (int j = 0; j < SPT; ++j)
#pragma unroll
    for (int q = 0; q < G; ++q) acc[j][q] = make_float2(0.f, 0.f);

  const long long plane = (long long)n0 * n1;
  const long long vol = plane * n2;
  const float2* up = A.u + (long long)b * A.u_bs + (long long)x2 * plane;
  const float2* sp = A.maps + (long long)(cvalid ? c : 0) * vol + (long long)x2 * plane;
  auto lidx = [&](int k1, int m2) { return g * (R1 * XS) + k1 * XS + m2; };
  __syncthreads();

  // rows are streamed with the next row's map/image values prefetched into
  // registers while the current row is transformed (loads are unconditional:
  // a workgroup's coils past ncoils read coil 0's map; their accumulators are
  // never stored and no other coil reads them)
  float2 sv[R1], uv[R1];
  if (t < R2) {
#pragma unroll
    for (int m1 = 0; m1 < R1; ++m1) {
      sv[m1] = sp[R2 * m1 + t];
      uv[m1] = up[R2 * m1 + t];
    }
  }
  // the thread's row-FFT twiddles exp(-2 pi i t k1 / NR), fixed over the rows
  float2 tw1[R1];
#pragma unroll
  for (int k1 = 1; k1 < R1; ++k1) tw1[k1] = twr[(t * k1) % NR];
  // twiddle index (k0 * x0) mod n0 of each sample, advanced row by row
  int tix[SPT];
#pragma unroll
  for (int j = 0; j < SPT; ++j) tix[j] = 0;
  for (int x0 = 0; x0 < n0; ++x0) {
    if (t < R2) {
      float2 a[R1];
#pragma unroll
      for (int m1 = 0; m1 < R1; ++m1) a[m1] = c_mul(sv[m1], uv[m1]);
      {  // unconditional (the last row reloads itself), so the loads land in place
        const int ro = min(x0 + 1, n0 - 1) * n1 + t;
#pragma unroll
        for (int m1 = 0; m1 < R1; ++m1) {
          sv[m1] = sp[ro + R2 * m1];
          uv[m1] = up[ro + R2 * m1];
        }
      }
      RDFT<R1, -1>::run(a);
#pragma unroll
      for (int k1 = 1; k1 < R1; ++k1) a[k1] = c_mul(a[k1], tw1[k1]);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) xch[lidx(k1, t)] = a[k1];
    }
    __syncthreads();
    if (t < R1) {
      float2 X[R2];
#pragma unroll
      for (int m2 = 0; m2 < R2; m2 += 2) {
        const float4 q = *reinterpret_cast<const float4*>(&xch[lidx(t, m2)]);
        X[m2] = make_float2(q.x, q.y);
        X[m2 + 1] = make_float2(q.z, q.w);
      }
      RDFT<R2, -1>::run(X);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) zrow[(t + R1 * k2) * ZS + g] = X[k2];
    }
    // the samples' column twiddles do not depend on this row's spectra
    float2 wj[SPT];
#pragma unroll
    for (int j = 0; j < SPT; ++j) wj[j] = twc[tix[j]];
    __syncthreads();
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
      if (FULL || mv[j]) {
#pragma unroll
        for (int q = 0; q < G; q += 2) {
          const float4 z = *reinterpret_cast<const float4*>(&zrow[mk1[j] * ZS + q]);
          acc[j][q] = c_fma(wj[j], make_float2(z.x, z.y), acc[j][q]);
          acc[j][q + 1] = c_fma(wj[j], make_float2(z.z, z.w), acc[j][q + 1]);
        }
      }
      tix[j] += mk0[j];
      tix[j] -= (tix[j] >= n0) ? n0 : 0;
    }
    __builtin_amdgcn_s_setprio(0);
  }

  const float scale = rsqrtf((float)n0 * (float)n1);
  double local = 0.0;
  const long long ybase = (long long)off * n2 * A.ncoils;
#pragma unroll
  for (int j = 0; j < SPT; ++j) {
    if (!mv[j]) continue;
    const int w = chunk * (SPT * NT) + j * NT + tid;
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int cc = cg * G + q;
      if (cc >= A.ncoils) continue;
      const long long idx = ybase + ((long long)x2 * A.ncoils + cc) * cnt + w;
      const float2 val = c_scale(acc[j][q], scale);
      const float2 res = A.meas ? c_sub(val, A.meas[idx]) : val;
      if (A.y) A.y[idx] = A.store_raw ? val : res;
      local += (double)res.x * res.x + (double)res.y * res.y;
    }
  }
  if (A.partial) {
    const double s = block_sum_f64(local, red);
    if (tid == 0) A.partial[(long long)b * A.blocks_per_shot + lb / A.nbatch] = s;
  }
}

// (k*x) mod n for 0 <= k, x < n <= 512: a mask when n is a power of two,
// else one float-reciprocal quotient estimate and a single correction.
__device__ __forceinline__ int mulmod(int k, int x, int n, int mask, float inv_n) {
  const int p = (int)__umul24(k, x);  // k, x < 512: the product fits 24 bits
  if (mask >= 0) return p & mask;
  int q = (int)((float)p * inv_n);
  int r = p - q * n;
  r += (r < 0) ? n : 0;
  r -= (r >= n) ? n : 0;
  return r;
}

// Wave-wide maximum of a non-negative int, returned as a uniform value.
__device__ __forceinline__ int wave_max_uniform(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return __builtin_amdgcn_readfirstlane(v);
}

// Rows per workgroup of the general adjoint (k_encode_adj_rows): 16 (a
// wavefront holds four column phases of all rows), 8 for lines of 32 threads.
template <int NR>
struct AdjT {
  static constexpr int R1 = RowT<NR>::R1, R2 = RowT<NR>::R2, T = RowT<NR>::T;
  static constexpr int L = T >= 32 ? 8 : 16;
  static constexpr int NT = T * L;
  static constexpr int RS = R1 * (R2 + 1) + 1;  // exchange row stride (odd: rows on distinct banks)
  // LDS carve (bytes): column twiddles at 0, staged sample records, exchange,
  // row twiddles, column pointers
  static constexpr int OFF_REC = kMaxCol * 8;
  static constexpr int OFF_XCH = OFF_REC + (kAdjChunk + 1) * 16;
  static constexpr int OFF_TWR = OFF_XCH + L * RS * 8;
  static constexpr int OFF_CP = OFF_TWR + NR * 8;
  static constexpr int BYTES = OFF_CP + (NR + 1) * 4;
};

// Adjoint for lines of more than 16 threads.  One workgroup per (readout
// plane x2, block of L rows, shot); each row is one line of T threads (the
// line-pass layout) and the workgroup
// loops over the coils, accumulating the coil sum of its rows in registers:
//   Z_c(x0, k1) = sum_{k0 in column k1} e^{+2 pi i k0 x0/N0} Y_c(k0, k1)   (scatter)
//   z_c(x0, :)  = IFFT_1 Z_c(x0, :)                                         (row IFFT)
//   out(x0, :) += conj(S_c(x0, :)) z_c(x0, :)                              (coil sum)
// Thread (r, t) = (tid % L, tid / L): a wavefront holds all rows of 64/L
// column phases t, so the samples it scatters for column k1 = R2*m1 + t come
// from only 64/L columns; the per-m1 trip count is the wave's longest of
// them (a uniform loop without divergent exits) and shorter columns read a
// zero record.  A coil's samples are staged in LDS as 16-byte records
// (y, k0 pre-scaled for the twiddle address); the next coil's samples and map
// values are prefetched into registers while the current coil is transformed.
template <int NR, bool POW2>
__global__ void __launch_bounds__(AdjT<NR>::NT) k_encode_adj_rows(EncArgs A) {
  using AT = AdjT<NR>;
  constexpr int R1 = AT::R1, R2 = AT::R2, NT = AT::NT, L = AT::L, RS = AT::RS;
  constexpr int SPT = (kAdjChunk + NT - 1) / NT;  // staged samples per thread
  constexpr int ZS = kAdjChunk;                   // index of the zero record
  __shared__ __align__(16) char smem[AT::BYTES];
  float2* twc = (float2*)smem;  // at LDS offset 0: a byte index is an address
  float4* rec = (float4*)(smem + AT::OFF_REC);
  float2* xch = (float2*)(smem + AT::OFF_XCH);
  float2* twr = (float2*)(smem + AT::OFF_TWR);
  int* cp = (int*)(smem + AT::OFF_CP);

  const int n0 = A.n0, n1 = A.n1, n2 = A.n2;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int b = lb % A.nbatch;  // shot fastest (map rows shared through L2)
  const int x2 = (lb / A.nbatch) % n2;
  const int row0 = (lb / A.nbatch / n2) * L;
  const int tid = threadIdx.x;
  const int r = tid % L, t = tid / L;
  const int x0 = row0 + r;
  const bool rvalid = x0 < n0;
  const int xr = rvalid ? x0 : 0;
  const int cnt = A.cnt[b], off = A.off[b];
  const int C = A.ncoils;
  const int mask8 = 8 * n0 - 1;  // POW2: (8 k0 x) mod 8 n0 = 8 ((k0 x) mod n0)
  const float inv_n0 = 1.0f / (float)n0;

  fill_twiddles<NR, NT>(twr);
  for (int j = tid; j < n0; j += NT) {
    float s, co;
    const int jj = (2 * j > n0) ? j - n0 : j;
    sincospif(-2.0f * (float)jj / (float)n0, &s, &co);
    twc[j] = make_float2(co, s);
  }
  for (int j = tid; j <= n1; j += NT) cp[j] = A.colptr[(long long)b * (n1 + 1) + j];
  if (tid == 0) rec[ZS] = make_float4(0.f, 0.f, 0.f, 0.f);

  const long long plane = (long long)n0 * n1;
  const long long vol = plane * n2;
  const long long ybase = (long long)off * n2 * C + (long long)x2 * C * cnt;  // coil c at ybase + c*cnt
  const float scale = rsqrtf((float)n0 * (float)n1);
  const float2* srow = A.maps + (long long)x2 * plane + (long long)xr * n1;  // coil c at srow + c*vol
  auto lidx = [&](int k1, int m2) { return r * RS + k1 * (R2 + 1) + m2; };

  float2 acc[R2];
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) acc[k2] = make_float2(0.f, 0.f);

  for (int s0 = 0; s0 < cnt; s0 += kAdjChunk) {
    const int ns = min(kAdjChunk, cnt - s0);
    __syncthreads();  // readers of the previous chunk's tables are done
    for (int j = tid; j < ns; j += NT) {
      const int k0 = A.sk0[off + s0 + j];
      rec[j].z = __int_as_float(POW2 ? 8 * k0 : k0);
    }
    // prefetch coil 0's samples and map row
    float2 ypre[SPT];
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
      const int w = tid + j * NT;
      ypre[j] = A.y[ybase + s0 + (w < ns ? w : 0)];
    }
    float2 sv[R2];
    if (t < R1) {
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) sv[k2] = srow[t + R1 * k2];
    }
    __syncthreads();  // cp visible
    // the thread's record ranges (byte offsets) within this chunk for its
    // columns k1 = R2*m1 + t (the same for every coil) and the wave's longest
    // column per m1
    int lo16[R1], wct[R1], nmax[R1];
#pragma unroll
    for (int m1 = 0; m1 < R1; ++m1) {
      int lo = 0, hi = 0;
      if (t < R2) {
        const int k1 = R2 * m1 + t;
        lo = max(cp[k1], s0) - s0;
        hi = max(min(cp[k1 + 1], s0 + ns) - s0, lo);
      }
      lo16[m1] = 16 * lo;
      wct[m1] = hi - lo;
      nmax[m1] = wave_max_uniform(hi - lo);
    }
    const char* recb = smem + AT::OFF_REC;
    for (int c = 0; c < C; ++c) {
      // every reader of the previous coil's records passed that coil's
      // exchange barrier, so one record buffer suffices
#pragma unroll
      for (int j = 0; j < SPT; ++j) {
        const int w = tid + j * NT;
        if (w < ns) *(float2*)&rec[w] = ypre[j];
      }
      if (c + 1 < C) {  // next coil's samples, in flight during this coil
#pragma unroll
        for (int j = 0; j < SPT; ++j) {
          const int w = tid + j * NT;
          ypre[j] = A.y[ybase + (long long)(c + 1) * cnt + s0 + (w < ns ? w : 0)];
        }
      }
      __syncthreads();
      // step 1: scatter this row's spectrum, length-R1 inverse DFTs
      if (t < R2) {
        float2 a[R1];
        // z += e^{+2 pi i k0 x0/n0} y for record j of the column (the zero
        // record past the column's end)
        auto term = [&](float2& z, int j, int ct, int lo) {
          const int o = j < ct ? lo + 16 * j : 16 * ZS;
          const float4 q = *(const float4*)(recb + o);
          const int k = __float_as_int(q.z);
          float2 w;
          if (POW2)
            w = *(const float2*)(smem + (int)(__umul24(k, xr) & mask8));
          else
            w = twc[mulmod(k, xr, n0, -1, inv_n0)];
          z = c_fmac(w, make_float2(q.x, q.y), z);  // twc holds e^{-2 pi i j/n0}: conjugate
        };
#pragma unroll
        for (int m1 = 0; m1 < R1; ++m1) {
          // two records per step: two independent load chains in flight
          float2 z0 = make_float2(0.f, 0.f), z1 = make_float2(0.f, 0.f);
          const int lo = lo16[m1], ct = wct[m1], nm = nmax[m1];
          int j = 0;
          for (; j + 1 < nm; j += 2) {
            term(z0, j, ct, lo);
            term(z1, j + 1, ct, lo);
          }
          if (j < nm) term(z0, j, ct, lo);
          a[m1] = c_add(z0, z1);
        }
        RDFT<R1, +1>::run(a);
#pragma unroll
        for (int k1 = 1; k1 < R1; ++k1) a[k1] = c_mulc(a[k1], twr[(t * k1) % NR]);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) xch[lidx(k1, t)] = a[k1];
      }
      __syncthreads();
      // step 2: length-R2 inverse DFTs, times conj(S_c), accumulated
      if (t < R1) {
        float2 X[R2];
#pragma unroll
        for (int m2 = 0; m2 < R2; ++m2) X[m2] = xch[lidx(t, m2)];
        RDFT<R2, +1>::run(X);
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) acc[k2] = c_add(acc[k2], c_mulc(X[k2], sv[k2]));
        if (c + 1 < C) {
          const float2* sn = srow + (long long)(c + 1) * vol;
#pragma unroll
          for (int k2 = 0; k2 < R2; ++k2) sv[k2] = sn[t + R1 * k2];
        }
      }
    }
  }
  if (t < R1 && rvalid) {
    float2* op = A.out + (long long)b * A.out_bs + (long long)x2 * plane + (long long)x0 * n1;
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) op[t + R1 * k2] = c_scale(acc[k2], scale);
  }
}

// Adjoint for lines of at most 16 threads (N1 <= 256): one workgroup per
// (readout plane x2, block of 64 rows, shot), looping over the coils.  In the
// scatter a wavefront is one column phase t and its 64 lanes are the rows, so
// every lane walks the same samples (the records of columns k1 = R2*m1 + t,
// read at wave-uniform LDS addresses with scalar loop bounds) and only the
// twiddle e^{+2 pi i k0 x0/N0} differs per lane; the length-R1 DFTs over m1
// are lane-local.  The exchange then hands each row to T threads (the
// line-pass layout, consecutive x1 across lanes) for the length-R2 DFTs, the
// conj(S_c) product and the coil sum in registers.
//
// Row twiddles e^{-2 pi i t k1/NR} (t < R2 the column phase, k1 < R1),
// phase-major: a wavefront's R1 twiddles are one contiguous run of 2 R1
// floats, fetched as a few wide scalar loads under one wait (indexing the
// length-NR table by (t k1) mod NR costs one dependent scalar load each).
template <int NR, int R1, int R2>
struct PhaseTw {
  alignas(64) float v[2 * R1 * R2];
  constexpr PhaseTw() : v{} {
    for (int t = 0; t < R2; ++t)
      for (int k = 0; k < R1; ++k) {
        const CS w = ce_cs_2pi((long)t * k, NR);
        v[2 * (t * R1 + k)] = (float)w.c;
        v[2 * (t * R1 + k) + 1] = -(float)w.s;
      }
  }
};
template <int NR, int R1, int R2>
__device__ constexpr PhaseTw<NR, R1, R2> kPhTw{};

template <int NR>
struct AdjW {
  static constexpr int R1 = RowT<NR>::R1, R2 = RowT<NR>::R2, T = RowT<NR>::T;
  static constexpr int L = 64;
  static constexpr int NT = T * L;
  static constexpr int RS = R1 * (R2 + 1) + 1;  // exchange row stride (odd)
  static constexpr int TWB = 8192;               // column twiddle tables at LDS offset 0
  static constexpr int OFF_REC = TWB;
  static constexpr int OFF_XCH = OFF_REC + kAdjChunk * 16;
  static constexpr int OFF_CP = OFF_XCH + L * RS * 8;
  static constexpr int BYTES = OFF_CP + (NR + 1) * 4;
  static_assert(kMaxCol * 8 <= TWB, "one twiddle table of the longest column");
};

// The scatter's cost is the latency of its dependent chain (record read,
// twiddle read, complex FMA) per sample.  The twiddle index is (k0 x0 mod
// N0): with k0 = 2^a k' the 64 rows of a wavefront touch entries spaced 2^a
// apart, a bank conflict of that order.  So the twiddle area holds one table
// per 2-adic order a up to that of N0 (entry j of table a is
// e^{-2 pi i (j 2^a mod N0)/N0}; (2^a (k' x0 mod N0)) mod N0 = k0 x0 mod N0
// for any N0) and a record carries k' and the byte base of table a:
// consecutive rows then read entries of an odd multiplier, spread over the
// banks.  Orders beyond the last table use it with an even k'.
template <int NR, bool POW2>
__global__ void __launch_bounds__(AdjW<NR>::NT) k_encode_adj_wave(EncArgs A) {
  using AW = AdjW<NR>;
  constexpr int R1 = AW::R1, R2 = AW::R2, T = AW::T, NT = AW::NT, L = AW::L, RS = AW::RS;
  constexpr int SPT = (kAdjChunk + NT - 1) / NT;  // staged samples per thread
  __shared__ __align__(16) char smem[AW::BYTES];
  float2* twc = (float2*)smem;
  float4* rec = (float4*)(smem + AW::OFF_REC);
  float2* xch = (float2*)(smem + AW::OFF_XCH);
  int* cp = (int*)(smem + AW::OFF_CP);

  const int n0 = A.n0, n1 = A.n1, n2 = A.n2;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int b = lb % A.nbatch;  // shot fastest (map rows shared through L2)
  const int x2 = (lb / A.nbatch) % n2;
  const int row0 = (lb / A.nbatch / n2) * L;
  const int tid = threadIdx.x;
  // scatter layout: lane = row, wavefront = column phase
  const int r = tid % L, t = tid / L;
  const int xr = (row0 + r < n0) ? row0 + r : 0;
  // line layout: T consecutive threads per row
  const int r2 = tid / T, t2 = tid % T;
  const int x0b = row0 + r2;
  const bool rvalid2 = x0b < n0;
  const int xr2 = rvalid2 ? x0b : 0;
  const int cnt = A.cnt[b], off = A.off[b];
  const int C = A.ncoils;
  const int mask8 = 8 * n0 - 1;
  const float inv_n0 = 1.0f / (float)n0;

  // twiddle tables: 2-adic orders 0 .. nsub-1 (up to the order of n0 itself)
  const int nsub = min(AW::TWB / (8 * n0), __builtin_ctz(n0) + 1);
  for (int e = tid; e < nsub * n0; e += NT) {
    const int a = e / n0, j = e % n0;
    const int jm = (int)(((long long)j << a) % n0);
    float s, co;
    const int jj = (2 * jm > n0) ? jm - n0 : jm;
    sincospif(-2.0f * (float)jj / (float)n0, &s, &co);
    twc[e] = make_float2(co, s);
  }
  for (int j = tid; j <= n1; j += NT) cp[j] = A.colptr[(long long)b * (n1 + 1) + j];

  const long long plane = (long long)n0 * n1;
  const long long vol = plane * n2;
  const long long ybase = (long long)off * n2 * C + (long long)x2 * C * cnt;  // coil c at ybase + c*cnt
  const float scale = rsqrtf((float)n0 * (float)n1);
  const float2* srow = A.maps + (long long)x2 * plane + (long long)xr2 * n1;  // coil c at srow + c*vol
  const char* recb = smem + AW::OFF_REC;

  float2 acc[R2];
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) acc[k2] = make_float2(0.f, 0.f);

  for (int s0 = 0; s0 < cnt; s0 += kAdjChunk) {
    const int ns = min(kAdjChunk, cnt - s0);
    __syncthreads();  // readers of the previous chunk's tables are done
    for (int j = tid; j < ns; j += NT) {
      const int k0 = A.sk0[off + s0 + j];
      const int a = k0 ? min(__builtin_ctz(k0), nsub - 1) : 0;
      rec[j].z = __int_as_float(POW2 ? 8 * (k0 >> a) : (k0 >> a));
      rec[j].w = __int_as_float(a * 8 * n0);  // byte base of table a
    }
    float2 ypre[SPT];
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
      const int w = tid + j * NT;
      ypre[j] = A.y[ybase + s0 + (w < ns ? w : 0)];
    }
    __syncthreads();  // cp and the k0 fields visible
    // the wavefront's record range (byte offsets) within this chunk for each
    // of its columns k1 = R2*m1 + t: wave-uniform
    int lo16[R1], hi16[R1];
#pragma unroll
    for (int m1 = 0; m1 < R1; ++m1) {
      int lo = 0, hi = 0;
      if (t < R2) {
        const int k1 = R2 * m1 + t;
        lo = max(cp[k1], s0) - s0;
        hi = max(min(cp[k1 + 1], s0 + ns) - s0, lo);
      }
      lo16[m1] = __builtin_amdgcn_readfirstlane(16 * lo);
      hi16[m1] = __builtin_amdgcn_readfirstlane(16 * hi);
    }
    for (int c = 0; c < C; ++c) {
      // every reader of the previous coil's records passed that coil's
      // exchange barrier, so one record buffer suffices
#pragma unroll
      for (int j = 0; j < SPT; ++j) {
        const int w = tid + j * NT;
        if (w < ns) *(float2*)&rec[w] = ypre[j];
      }
      if (c + 1 < C) {  // next coil's samples, in flight during this coil
#pragma unroll
        for (int j = 0; j < SPT; ++j) {
          const int w = tid + j * NT;
          ypre[j] = A.y[ybase + (long long)(c + 1) * cnt + s0 + (w < ns ? w : 0)];
        }
      }
      __syncthreads();
      // step 1: scatter (lane = row), length-R1 inverse DFTs
      if (t < R2) {
        float2 a[R1];
        auto term = [&](float2& z, int o) {  // z += e^{+2 pi i k0 x0/n0} y of the record at byte o
          const float4 q = *(const float4*)(recb + o);
          const int k = __float_as_int(q.z);
          float2 w;
          if (POW2) {
            // table base | 8 ((k' x0) mod n0), the base a multiple of 8 n0, as one
            // v_and_or_b32 (the compiler emits a separate and and or here)
            int ad;
            asm("v_and_or_b32 %0, %1, %2, %3"
                : "=v"(ad)
                : "v"((int)__umul24(k, xr)), "s"(mask8), "v"(__float_as_int(q.w)));
            w = *(const float2*)(smem + ad);
          } else {
            w = *(const float2*)(smem + 8 * mulmod(k, xr, n0, -1, inv_n0) + __float_as_int(q.w));
          }
          z = c_fmac(w, make_float2(q.x, q.y), z);  // twc holds e^{-2 pi i j/n0}: conjugate
        };
        // one accumulator per column, zeroed once for all columns up front
        // (per-column zeroing inside the loop nest is rematerialised on both
        // sides of every loop guard)
#pragma unroll
        for (int m1 = 0; m1 < R1; ++m1) a[m1] = make_float2(0.f, 0.f);
#pragma unroll
        for (int m1 = 0; m1 < R1; ++m1) {
          int o = lo16[m1];
          const int oe = hi16[m1];
          for (; o + 16 < oe; o += 32) {
            term(a[m1], o);
            term(a[m1], o + 16);
          }
          if (o < oe) term(a[m1], o);
        }
        RDFT<R1, +1>::run(a);
        {  // row twiddles of the wavefront's column phase t
          const int tu = __builtin_amdgcn_readfirstlane(t);
          const float* tp = kPhTw<NR, R1, R2>.v + 2 * R1 * tu;
#pragma unroll
          for (int k1 = 1; k1 < R1; ++k1) a[k1] = c_mulc(a[k1], make_float2(tp[2 * k1], tp[2 * k1 + 1]));
        }
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) xch[r * RS + k1 * (R2 + 1) + t] = a[k1];
      }
      __syncthreads();
      // step 2 (line layout): length-R2 inverse DFTs, times conj(S_c),
      // accumulated; the map loads are issued first and consumed last (the
      // exchange reads and the DFT of the 16 wavefronts cover their latency)
      if (t2 < R1) {
        float2 sv[R2];
        const float2* sc = srow + (long long)c * vol;
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) sv[k2] = sc[t2 + R1 * k2];
        float2 X[R2];
#pragma unroll
        for (int m2 = 0; m2 < R2; ++m2) X[m2] = xch[r2 * RS + t2 * (R2 + 1) + m2];
        RDFT<R2, +1>::run(X);
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) acc[k2] = c_add(acc[k2], c_mulc(X[k2], sv[k2]));
      }
    }
  }
  if (t2 < R1 && rvalid2) {
    float2* op = A.out + (long long)b * A.out_bs + (long long)x2 * plane + (long long)x0b * n1;
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) op[t2 + R1 * k2] = c_scale(acc[k2], scale);
  }
}

// Fixed-order sum of per-workgroup partials: out[b] = sum_i partial[b*nper + i]
__global__ void k_sum_partials(const double* __restrict__ partial, int nper, double* __restrict__ out) {
  __shared__ double sh[16];
  double v = 0.0;
  const double* p = partial + (long long)blockIdx.x * nper;
  for (int i = threadIdx.x; i < nper; i += blockDim.x) v += p[i];
  const double s = block_sum_f64(v, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// y[c] = scale * x[(c + off) mod xdims] over 4-D (c, x0, x1, x2) index
// spaces with arbitrary element strides (ydims <= xdims): covers fftshift /
// ifftshift (off = n/2 or n - n/2), centred crops (off = n/2 - m/2) and
// layout changes.  Threads walk y in its memory order (ord[0] slowest).
struct PermArgs {
  const float2* x;
  float2* y;
  long long xs[4], ys[4];
  int ydims[4], xdims[4], off[4], ord[4];
  float scale;
  long long total;
};
__global__ void k_permute_shift(PermArgs P) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.total) return;
  long long rem = i;
  int co[4];
#pragma unroll
  for (int k = 3; k >= 0; --k) {
    const int a = P.ord[k];
    co[a] = (int)(rem % P.ydims[a]);
    rem /= P.ydims[a];
  }
  long long xo = 0, yo = 0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    int s = co[a] + P.off[a];
    if (s >= P.xdims[a]) s -= P.xdims[a];
    xo += (long long)s * P.xs[a];
    yo += (long long)co[a] * P.ys[a];
  }
  const float2 v = P.x[xo];
  P.y[yo] = make_float2(v.x * P.scale, v.y * P.scale);
}

// Sum over coils of |S_c|^2 (plane-major or any layout; maps coil-contiguous):
//   P = 1 / (ssq + eps), M = ssq > thr   (joint_recon.py:83-85)
__global__ void k_coil_ssq(const float2* __restrict__ maps, int ncoils, long long vol, float eps, float thr,
                           float* __restrict__ P, unsigned char* __restrict__ M) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vol) return;
  float s = 0.f;
  for (int c = 0; c < ncoils; ++c) {
    const float2 v = maps[(long long)c * vol + i];
    s = fmaf(v.x, v.x, fmaf(v.y, v.y, s));
  }
  if (P) P[i] = 1.0f / (s + eps);
  if (M) M[i] = s > thr ? 1 : 0;
}

// Ymeas[(off*n2*C) + (x2*C + c)*cnt + w] = scale * ksp_h[c][k0][k1][x2]
// (ksp_h in reference C order, after the readout inverse FFT).
__global__ void k_gather_hybrid(const float2* __restrict__ ksp, int ncoils, int n0, int n1, int n2,
                                const int* __restrict__ sk0, const int* __restrict__ sk1, const int* __restrict__ off,
                                const int* __restrict__ cnt, int nshots, long long total_samples, float scale,
                                float2* __restrict__ ymeas) {
  // one thread per (global sample, coil); loop over x2 (reads contiguous in x2)
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total_samples * ncoils) return;
  const int c = (int)(i % ncoils);
  const long long gw = i / ncoils;
  // locate the shot of this sample (shots are contiguous, off increasing)
  int lo = 0, hi = nshots - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((long long)off[mid] <= gw) lo = mid; else hi = mid - 1;
  }
  const int s = lo;
  const int w = (int)(gw - off[s]);
  const long long src = (((long long)c * n0 + sk0[gw]) * n1 + sk1[gw]) * n2;
  const long long dbase = (long long)off[s] * n2 * ncoils;
  for (int x2 = 0; x2 < n2; ++x2) {
    const float2 v = ksp[src + x2];
    ymeas[dbase + ((long long)x2 * ncoils + c) * cnt[s] + w] = make_float2(v.x * scale, v.y * scale);
  }
}

// ksp_h[c][k0][k1][x2] += Y[...]: inverse of k_gather_hybrid (atomic adds, so
// shots may share lines).
__global__ void k_scatter_hybrid(const float2* __restrict__ y, int ncoils, int n0, int n1, int n2,
                                 const int* __restrict__ sk0, const int* __restrict__ sk1, const int* __restrict__ off,
                                 const int* __restrict__ cnt, int nshots, long long total_samples,
                                 float2* __restrict__ ksp) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total_samples * ncoils) return;
  const int c = (int)(i % ncoils);
  const long long gw = i / ncoils;
  int lo = 0, hi = nshots - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((long long)off[mid] <= gw) lo = mid; else hi = mid - 1;
  }
  const int s = lo;
  const int w = (int)(gw - off[s]);
  const long long dst = (((long long)c * n0 + sk0[gw]) * n1 + sk1[gw]) * n2;
  const long long sbase = (long long)off[s] * n2 * ncoils;
  for (int x2 = 0; x2 < n2; ++x2) {
    const float2 v = y[sbase + ((long long)x2 * ncoils + c) * cnt[s] + w];
    atomicAdd(&ksp[dst + x2].x, v.x);
    atomicAdd(&ksp[dst + x2].y, v.y);
  }
}

// max |x| -> atomicMax on the float bit pattern (non-negative floats order as ints)
__global__ void k_maxabs(const float2* __restrict__ x, long long n, unsigned int* __restrict__ out) {
  float m = 0.f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float2 v = x[i];
    m = fmaxf(m, sqrtf(v.x * v.x + v.y * v.y));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_down(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

namespace {
template <int SPT, int G>
struct FwdLaunch {
  const EncArgs* a;
  dim3 grid;
  hipStream_t s;
  template <int NR>
  void operator()() {
    if (a->max_cnt % (SPT * FwdT<NR, G>::NT) == 0)
      hipLaunchKernelGGL((k_encode_fwd<NR, SPT, G, true>), grid, dim3(FwdT<NR, G>::NT), 0, s, *a);
    else
      hipLaunchKernelGGL((k_encode_fwd<NR, SPT, G, false>), grid, dim3(FwdT<NR, G>::NT), 0, s, *a);
  }
};
bool launch_fwd(int variant, int n1, const EncArgs* a, dim3 grid, hipStream_t s) {
  switch (variant) {
    case 0: { FwdLaunch<2, 16> f{a, grid, s}; return dispatch_line_length(n1, f); }
    case 1: { FwdLaunch<2, 8> f{a, grid, s}; return dispatch_line_length(n1, f); }
    case 2: { FwdLaunch<4, 8> f{a, grid, s}; return dispatch_line_length(n1, f); }
    case 3: { FwdLaunch<4, 4> f{a, grid, s}; return dispatch_line_length(n1, f); }
    default: return false;
  }
}
// lines of at most 16 threads use the wave-per-phase adjoint
template <int NR>
constexpr bool adj_wave() { return RowT<NR>::T <= 16; }
struct AdjLaunch {
  const EncArgs* a;
  dim3 grid;
  size_t shmem;
  hipStream_t s;
  template <int NR>
  void operator()() {
    const bool pow2 = (a->n0 & (a->n0 - 1)) == 0;
    if constexpr (adj_wave<NR>()) {
      if (pow2)
        hipLaunchKernelGGL((k_encode_adj_wave<NR, true>), grid, dim3(AdjW<NR>::NT), shmem, s, *a);
      else
        hipLaunchKernelGGL((k_encode_adj_wave<NR, false>), grid, dim3(AdjW<NR>::NT), shmem, s, *a);
    } else {
      if (pow2)
        hipLaunchKernelGGL((k_encode_adj_rows<NR, true>), grid, dim3(AdjT<NR>::NT), shmem, s, *a);
      else
        hipLaunchKernelGGL((k_encode_adj_rows<NR, false>), grid, dim3(AdjT<NR>::NT), shmem, s, *a);
    }
  }
};
struct AdjRows {
  int* rows;
  template <int NR>
  void operator()() { *rows = adj_wave<NR>() ? AdjW<NR>::L : AdjT<NR>::L; }
};
struct RowThreads {
  int* t;
  template <int NR>
  void operator()() { *t = RowT<NR>::T; }
};
// threads per row of N1 (0: unsupported length)
int threads_for_row(int n1) {
  int t = 0;
  RowThreads f{&t};
  dispatch_line_length(n1, f);
  return t;
}
// workgroups per shot of a forward-encoding variant
struct FwdGeom {
  int ncg, nchunk, blocks_per_shot;
};
FwdGeom fwd_geom(int variant, int n1, int n2, int ncoils, int max_cnt) {
  const FwdVariant v = kFwdVariants[variant];
  const int nt = threads_for_row(n1) * v.G;
  FwdGeom g{};
  g.ncg = (ncoils + v.G - 1) / v.G;
  g.nchunk = max_cnt > 0 ? (max_cnt + v.SPT * nt - 1) / (v.SPT * nt) : 1;
  g.blocks_per_shot = n2 * g.ncg * g.nchunk;
  return g;
}
// tuned variant per (N0, N1, N2, coils, largest shot)
std::mutex g_enc_mu;
std::map<std::tuple<int, int, int, int, int>, int> g_enc_tune;
int fwd_variant(const int* n, int ncoils, int max_cnt) {
  std::lock_guard<std::mutex> lk(g_enc_mu);
  const auto it = g_enc_tune.find(std::make_tuple(n[0], n[1], n[2], ncoils, max_cnt));
  return it == g_enc_tune.end() ? 0 : it->second;
}
}  // namespace

// Norm partials of me_encode_forward_static: a file-scope device array (each
// GPU holds its own copy until the process ends).  Host code never passes the
// symbol itself to a kernel: its device address, looked up once per device
// with hipGetSymbolAddress, is what the kernels receive.  Every entry a call
// reads (k_sum_partials) is written by that call's k_encode_fwd first, so
// nothing needs clearing; calls that overlap on one GPU share the array and
// race.
constexpr long long kStaticPartials = ME_ENCODE_STATIC_PARTIALS;
__device__ __align__(256) double scratch[kStaticPartials];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the device that runs `stream` (which must be
// the current device).
int static_partials(hipStream_t stream, double** out) {
  int dev = -1;
  ME_TRY_HIP(hipGetDevice(&dev));
  if (stream) {
    int sdev = -1;
    ME_TRY_HIP(hipStreamGetDevice(stream, &sdev));
    ME_REQUIRE(sdev == dev, "encode_forward_static: the stream belongs to device " + std::to_string(sdev) +
                                ", not the current device " + std::to_string(dev));
  }
  ME_REQUIRE(dev >= 0 && dev < kMaxDevices, "encode_forward_static: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    ME_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ME_REQUIRE(p != nullptr, "encode_forward_static: no device address for the scratch array");
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

int encode_forward_impl(const void* u, long long u_bstride, const void* maps, int ncoils, int nbatch, const int* n,
                        const int* samp_k0, const int* samp_k1, const int* shot_off, const int* shot_cnt, int max_cnt,
                        void* y, const void* y_meas, double* partials, double* norms, int flags, hipStream_t stream) {
  ME_REQUIRE(u && maps && n && samp_k0 && samp_k1 && shot_off && shot_cnt, "encode_forward: null argument");
  ME_REQUIRE(nbatch >= 1 && nbatch <= 65535 && ncoils >= 1, "encode_forward: bad batch/coil count");
  ME_REQUIRE(n[0] <= kMaxCol, "encode_forward: N0 above 512");
  ME_REQUIRE(!norms || partials, "encode_forward: norms need a partials workspace");
  ME_REQUIRE(threads_for_row(n[1]) > 0, "encode_forward: unsupported N1 " + std::to_string(n[1]));
  const int variant = fwd_variant(n, ncoils, max_cnt);
  const FwdGeom fg = fwd_geom(variant, n[1], n[2], ncoils, max_cnt);
  EncArgs A{};
  A.u = (const float2*)u;
  A.u_bs = u_bstride;
  A.maps = (const float2*)maps;
  A.ncoils = ncoils;
  A.n0 = n[0];
  A.n1 = n[1];
  A.n2 = n[2];
  A.sk0 = samp_k0;
  A.sk1 = samp_k1;
  A.off = shot_off;
  A.cnt = shot_cnt;
  A.y = (float2*)y;
  A.meas = (const float2*)y_meas;
  A.partial = partials;
  A.store_raw = (flags & 1) ? 1 : 0;
  A.ncg = fg.ncg;
  A.nchunk = fg.nchunk;
  A.blocks_per_shot = fg.blocks_per_shot;
  A.max_cnt = max_cnt;
  A.nbatch = nbatch;
  ME_REQUIRE((long long)A.blocks_per_shot * nbatch < (1LL << 31), "encode_forward: grid too large");
  dim3 grid((unsigned)(A.blocks_per_shot * nbatch), 1u);
  ME_REQUIRE(launch_fwd(variant, n[1], &A, grid, stream), "encode_forward: unsupported N1");
  ME_TRY_HIP(hipGetLastError());
  if (norms) {
    hipLaunchKernelGGL(k_sum_partials, dim3(nbatch), dim3(256), 0, stream, partials, A.blocks_per_shot, norms);
    ME_TRY_HIP(hipGetLastError());
  }
  return 0;
}

}  // namespace me

using namespace me;

extern "C" {

long long me_encode_partials_size(int n2, int ncoils, int max_cnt, int n1) {
  if (threads_for_row(n1) == 0 || ncoils <= 0 || n2 <= 0) return 0;
  long long most = 0;  // enough for whichever variant is tuned in
  for (int v = 0; v < kNumFwdVariants; ++v) most = std::max(most, (long long)fwd_geom(v, n1, n2, ncoils, max_cnt).blocks_per_shot);
  return most;
}

int me_encode_tuning_set(const int* n, int ncoils, int max_cnt, int variant) {
  ME_REQUIRE(n && ncoils >= 1, "encode_tuning_set: bad argument");
  ME_REQUIRE(variant >= 0 && variant < kNumFwdVariants, "encode_tuning_set: variant must be 0, 1, 2 or 3");
  std::lock_guard<std::mutex> lk(g_enc_mu);
  const auto key = std::make_tuple(n[0], n[1], n[2], ncoils, max_cnt);
  if (variant == 0)
    g_enc_tune.erase(key);
  else
    g_enc_tune[key] = variant;
  return 0;
}

int me_encode_tuning_get(const int* n, int ncoils, int max_cnt) { return n ? fwd_variant(n, ncoils, max_cnt) : -1; }

void me_encode_tuning_clear(void) {
  std::lock_guard<std::mutex> lk(g_enc_mu);
  g_enc_tune.clear();
}

int me_tune_encode_forward(const void* u, long long u_bstride, const void* maps, int ncoils, int nbatch, const int* n,
                           const int* samp_k0, const int* samp_k1, const int* shot_off, const int* shot_cnt,
                           int max_cnt, void* y, int reps, int* chosen, float* times_ms, hipStream_t stream) {
  ME_REQUIRE(u && maps && n && y && reps >= 1, "tune_encode_forward: bad argument");
  float best = -1.f;
  int pick = 0;
  hipEvent_t e0, e1;
  ME_TRY_HIP(hipEventCreate(&e0));
  ME_TRY_HIP(hipEventCreate(&e1));
  int rc = 0;
  for (int v = 0; v < kNumFwdVariants && rc == 0; ++v) {
    rc = me_encode_tuning_set(n, ncoils, max_cnt, v);
    for (int r = 0; r <= reps && rc == 0; ++r) {  // the first launch is untimed
      if (r == 1 && hipEventRecord(e0, stream) != hipSuccess) rc = 1;
      if (rc == 0)
        rc = encode_forward_impl(u, u_bstride, maps, ncoils, nbatch, n, samp_k0, samp_k1, shot_off, shot_cnt,
                                 max_cnt, y, nullptr, nullptr, nullptr, 0, stream);
    }
    float ms = 0.f;
    if (rc == 0 && (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = 1;
    ms /= (float)reps;
    if (times_ms) times_ms[v] = ms;
    if (rc == 0 && (best < 0.f || ms < best)) {
      best = ms;
      pick = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != 0) {
    me_encode_tuning_set(n, ncoils, max_cnt, 0);
    if (rc == 1) ::me::set_error("tune_encode_forward: HIP event timing failed");
    return rc;
  }
  if (chosen) *chosen = pick;
  return me_encode_tuning_set(n, ncoils, max_cnt, pick);
}

int me_encode_forward(const void* u, long long u_bstride, const void* maps, int ncoils, int nbatch, const int* n,
                      const int* samp_k0, const int* samp_k1, const int* shot_off, const int* shot_cnt, int max_cnt,
                      void* y, const void* y_meas, double* partials, double* norms, int flags, hipStream_t stream) {
  return encode_forward_impl(u, u_bstride, maps, ncoils, nbatch, n, samp_k0, samp_k1, shot_off, shot_cnt, max_cnt, y,
                             y_meas, partials, norms, flags, stream);
}

int me_encode_forward_static(const void* u, long long u_bstride, const void* maps, int ncoils, int nbatch,
                             const int* n, const int* samp_k0, const int* samp_k1, const int* shot_off,
                             const int* shot_cnt, int max_cnt, void* y, const void* y_meas, double* norms, int flags,
                             hipStream_t stream) {
  double* partials = nullptr;
  if (norms) {
    ME_REQUIRE(n && nbatch >= 1, "encode_forward_static: bad argument");
    ME_REQUIRE(threads_for_row(n[1]) > 0, "encode_forward_static: unsupported N1 " + std::to_string(n[1]));
    const long long need =
        (long long)fwd_geom(fwd_variant(n, ncoils, max_cnt), n[1], n[2], ncoils, max_cnt).blocks_per_shot * nbatch;
    ME_REQUIRE(need <= kStaticPartials,
               "encode_forward_static: " + std::to_string(need) + " FP64 norm partials needed, the static scratch "
               "holds " + std::to_string(kStaticPartials) + " (ME_ENCODE_STATIC_PARTIALS); use me_encode_forward "
               "with a workspace or fewer shots per call");
    const int rc = static_partials(stream, &partials);
    if (rc) return rc;
  }
  return encode_forward_impl(u, u_bstride, maps, ncoils, nbatch, n, samp_k0, samp_k1, shot_off, shot_cnt, max_cnt, y,
                             y_meas, partials, norms, flags, stream);
}

int me_encode_coil_groups(int ncoils) { return ncoils >= 1 ? 1 : 0; }

int me_encode_adjoint(const void* y, const void* maps, int ncoils, int nbatch, const int* n, const int* samp_k0,
                      const int* shot_off, const int* shot_cnt, const int* colptr, void* out, long long out_bstride,
                      hipStream_t stream) {
  ME_REQUIRE(y && maps && n && samp_k0 && shot_off && shot_cnt && colptr && out, "encode_adjoint: null argument");
  ME_REQUIRE(nbatch >= 1 && nbatch <= 65535 && ncoils >= 1, "encode_adjoint: bad batch/coil count");
  ME_REQUIRE(n[0] <= kMaxCol, "encode_adjoint: N0 above 512");
  ME_REQUIRE(threads_for_row(n[1]) > 0, "encode_adjoint: unsupported N1 " + std::to_string(n[1]));
  EncArgs A{};
  A.y = (float2*)y;
  A.maps = (const float2*)maps;
  A.ncoils = ncoils;
  A.n0 = n[0];
  A.n1 = n[1];
  A.n2 = n[2];
  A.sk0 = samp_k0;
  A.off = shot_off;
  A.cnt = shot_cnt;
  A.colptr = colptr;
  A.out = (float2*)out;
  A.out_bs = out_bstride;
  A.ncg = 1;
  A.nbatch = nbatch;
  int rows = 0;
  AdjRows fr{&rows};
  dispatch_line_length(n[1], fr);
  const int nrb = (n[0] + rows - 1) / rows;
  ME_REQUIRE((long long)n[2] * nrb * nbatch < (1LL << 31), "encode_adjoint: grid too large");
  dim3 grid((unsigned)(n[2] * nrb * nbatch), 1u);
  AdjLaunch f{&A, grid, 0, stream};
  ME_REQUIRE(dispatch_line_length(n[1], f), "encode_adjoint: unsupported N1");
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_permute_shift(const void* x, const long long* x_strides, const int* x_dims, void* y,
                     const long long* y_strides, const int* y_dims, const int* offset, float scale,
                     hipStream_t stream) {
  ME_REQUIRE(x && y && x_strides && y_strides && x_dims && y_dims && offset, "permute_shift: null argument");
  PermArgs P{};
  P.x = (const float2*)x;
  P.y = (float2*)y;
  P.scale = scale;
  P.total = 1;
  for (int a = 0; a < 4; ++a) {
    ME_REQUIRE(y_dims[a] >= 1 && x_dims[a] >= y_dims[a], "permute_shift: output extent exceeds input extent");
    P.xs[a] = x_strides[a];
    P.ys[a] = y_strides[a];
    P.ydims[a] = y_dims[a];
    P.xdims[a] = x_dims[a];
    P.off[a] = ((offset[a] % x_dims[a]) + x_dims[a]) % x_dims[a];
    P.total *= y_dims[a];
    P.ord[a] = a;
  }
  for (int i = 0; i < 4; ++i)  // axes by decreasing output stride (slowest first)
    for (int j = i + 1; j < 4; ++j)
      if (P.ys[P.ord[j]] > P.ys[P.ord[i]]) {
        const int tmp = P.ord[i];
        P.ord[i] = P.ord[j];
        P.ord[j] = tmp;
      }
  const unsigned blocks = (unsigned)((P.total + 255) / 256);
  hipLaunchKernelGGL(k_permute_shift, dim3(blocks), dim3(256), 0, stream, P);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_coil_ssq(const void* maps, int ncoils, long long vol, float eps, float thr, float* P, unsigned char* M,
                hipStream_t stream) {
  ME_REQUIRE(maps && ncoils >= 1 && vol >= 1, "coil_ssq: bad argument");
  hipLaunchKernelGGL(k_coil_ssq, dim3((unsigned)((vol + 255) / 256)), dim3(256), 0, stream, (const float2*)maps,
                     ncoils, vol, eps, thr, P, M);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_gather_hybrid(const void* ksp_h, int ncoils, const int* n, const int* samp_k0, const int* samp_k1,
                     const int* shot_off, const int* shot_cnt, int nshots, long long total_samples, float scale,
                     void* ymeas, hipStream_t stream) {
  ME_REQUIRE(ksp_h && n && samp_k0 && samp_k1 && shot_off && shot_cnt && ymeas, "gather_hybrid: null argument");
  const long long work = total_samples * ncoils;
  if (work == 0) return 0;
  hipLaunchKernelGGL(k_gather_hybrid, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, stream,
                     (const float2*)ksp_h, ncoils, n[0], n[1], n[2], samp_k0, samp_k1, shot_off, shot_cnt, nshots,
                     total_samples, scale, (float2*)ymeas);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_scatter_hybrid(const void* y, int ncoils, const int* n, const int* samp_k0, const int* samp_k1,
                      const int* shot_off, const int* shot_cnt, int nshots, long long total_samples, void* ksp_h,
                      hipStream_t stream) {
  ME_REQUIRE(y && n && samp_k0 && samp_k1 && shot_off && shot_cnt && ksp_h, "scatter_hybrid: null argument");
  const long long work = total_samples * ncoils;
  if (work == 0) return 0;
  hipLaunchKernelGGL(k_scatter_hybrid, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, stream, (const float2*)y,
                     ncoils, n[0], n[1], n[2], samp_k0, samp_k1, shot_off, shot_cnt, nshots, total_samples,
                     (float2*)ksp_h);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_max_abs(const void* x, long long n, unsigned int* out_bits, hipStream_t stream) {
  ME_REQUIRE(x && out_bits, "max_abs: null argument");
  hipLaunchKernelGGL(k_maxabs, dim3(1024), dim3(256), 0, stream, (const float2*)x, n, out_bits);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
