// CG-SENSE vector steps and the Levenberg-Marquardt motion update.
//
// CG (image_estimation.py:130-165): all scalars live on the device in an FP64
// state vector so no step needs a host synchronisation; inner products are
// accumulated in FP64 (wavefront shuffles, one value per workgroup, then a
// fixed-order sum in a single workgroup), which keeps results deterministic.
// The preconditioned residual z = P r is never stored: it is recomputed
// where used.  The tolerance test of the reference (break after the x/r
// update when sqrt|<r, z>| < tol, image_estimation.py:156-157) sets a device
// flag that turns every later update into a no-op.
//
// LM (transform_estimation.py:39-150): per shot, e0 = ||r||^2, g = Re J^H r
// and the Gram matrix Re J^H J over the gathered samples (FP64), the damped
// 6x6 solve in FP64 (Gaussian elimination with partial pivoting), the
// candidate theta - delta, accept/reject with the reference's damping rule,
// and the mean re-registration of the parameters.
#include "me_device.h"
#include "me_internal.h"
#include "../../include/motion_estimation_amd.h"

namespace me {

constexpr int kVecBlocks = 1024;
constexpr int kDotBatchBlocks = 64;  // workgroups per batch entry of the batched dot
constexpr int kVecThreads = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Sum over the workgroup; the result is valid in thread 0.
template <int NT>
__device__ __forceinline__ double wg_sum(double v, double* sh) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NT / 64; ++i) s += sh[i];
  return s;
}

// Fixed-order sum of nblk partials by one workgroup; valid in thread 0.
__device__ __forceinline__ double sum_partials(const double* p, int nblk, double* sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kVecThreads) v += p[i];
  return wg_sum<kVecThreads>(v, sh);
}

// ---------------------------------------------------------------- dot
__global__ void __launch_bounds__(kVecThreads) k_dot_re(const float2* __restrict__ a, const float2* __restrict__ b,
                                                        const float* __restrict__ w, long long n,
                                                        double* __restrict__ partial) {
  __shared__ double sh[kVecThreads / 64];
  double v = 0.0;
  for (long long i = (long long)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kVecThreads) {
    const float2 x = a[i], y = b[i];
    double d = (double)x.x * y.x + (double)x.y * y.y;  // Re(conj(a) b)
    if (w) d *= (double)w[i];
    v += d;
  }
  const double s = wg_sum<kVecThreads>(v, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kVecThreads) k_finalize_sum(const double* __restrict__ partial, int nblk,
                                                              double* __restrict__ out) {
  __shared__ double sh[kVecThreads / 64];
  const double s = sum_partials(partial, nblk, sh);
  if (threadIdx.x == 0) *out = s;
}

// out[b] = Re <a_b, c_b> over n values per batch entry (grid (nblk, nbatch))
__global__ void __launch_bounds__(kVecThreads) k_dot_re_batched(const float2* __restrict__ a, long long a_bs,
                                                                const float2* __restrict__ c, long long c_bs,
                                                                long long n, double* __restrict__ partial) {
  __shared__ double sh[kVecThreads / 64];
  const float2* ab = a + (long long)blockIdx.y * a_bs;
  const float2* cb = c + (long long)blockIdx.y * c_bs;
  double v = 0.0;
  for (long long i = (long long)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kVecThreads) {
    const float2 x = ab[i], y = cb[i];
    v += (double)x.x * y.x + (double)x.y * y.y;
  }
  const double s = wg_sum<kVecThreads>(v, sh);
  if (threadIdx.x == 0) partial[(long long)blockIdx.y * gridDim.x + blockIdx.x] = s;
}

__global__ void __launch_bounds__(kVecThreads) k_finalize_batched(const double* __restrict__ partial, int nblk,
                                                                  double* __restrict__ out, long long out_stride) {
  __shared__ double sh[kVecThreads / 64];
  const double s = sum_partials(partial + (long long)blockIdx.x * nblk, nblk, sh);
  if (threadIdx.x == 0) out[(long long)blockIdx.x * out_stride] = s;
}

// out (+)= sum_b x_b over nbatch entries of n values (fixed order)
__global__ void k_sum_batch(const float2* __restrict__ x, long long x_bs, int nbatch, long long n,
                            float2* __restrict__ out, int accumulate) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    float2 acc = accumulate ? out[i] : make_float2(0.f, 0.f);
    for (int b = 0; b < nbatch; ++b) acc = c_add(acc, x[(long long)b * x_bs + i]);
    out[i] = acc;
  }
}

// y += a x (a real)
__global__ void k_axpy_real(float2* __restrict__ y, float a, const float2* __restrict__ x, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float2 v = x[i];
    float2 w = y[i];
    w.x = fmaf(a, v.x, w.x);
    w.y = fmaf(a, v.y, w.y);
    y[i] = w;
  }
}

// ---------------------------------------------------------------- CG
// CG state.  CG_RULE (set by the caller) selects the stopping rule:
// 0 = estimate_image_cg (image_estimation.py:156-157: stop after an update
//     when sqrt|<r, P r>| < tol);
// 1 = sigpy.alg.ConjugateGradient (as run by sigpy.mri.app.SenseRecon): stop
//     before any update when sqrt(<r, P r>) <= tol, when <p, A p> <= 0
//     (no update), and after an update when sqrt(<r, P r>) <= tol.
enum { CG_RSOLD = 0, CG_RSNEW, CG_PAP, CG_ALPHA, CG_BETA, CG_DONE, CG_ITERS, CG_RULE, CG_TOL, CG_NSTATE };

// r = b - Ap; p = P r; partial sums of P |r|^2
__global__ void __launch_bounds__(kVecThreads) k_cg_init(const float2* __restrict__ bv, const float2* __restrict__ Ap,
                                                         float2* __restrict__ r, float2* __restrict__ p,
                                                         const float* __restrict__ P, long long n,
                                                         double* __restrict__ partial) {
  __shared__ double sh[kVecThreads / 64];
  double v = 0.0;
  for (long long i = (long long)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kVecThreads) {
    const float2 ri = c_sub(bv[i], Ap[i]);
    const float pi = P ? P[i] : 1.0f;
    r[i] = ri;
    p[i] = c_scale(ri, pi);
    v += (double)pi * ((double)ri.x * ri.x + (double)ri.y * ri.y);
  }
  const double s = wg_sum<kVecThreads>(v, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// op 0: rsold = s, done = 0; op 1: pAp = s
__global__ void __launch_bounds__(kVecThreads) k_cg_finalize(const double* __restrict__ partial, int nblk,
                                                             double* __restrict__ st, int op, double tol) {
  __shared__ double sh[kVecThreads / 64];
  if (op == 2 && st[CG_DONE] != 0.0) return;
  const double s = sum_partials(partial, nblk, sh);
  if (threadIdx.x != 0) return;
  const bool sigpy = st[CG_RULE] == 1.0;
  if (op == 0) {
    st[CG_RSOLD] = s;
    st[CG_DONE] = (sigpy && sqrt(fabs(s)) <= st[CG_TOL]) ? 1.0 : 0.0;
    st[CG_ITERS] = 0.0;
  } else if (op == 1) {
    st[CG_PAP] = s;
    if (sigpy && s <= 0.0) st[CG_DONE] = 1.0;
  } else {  // op 2: after the x/r update (image_estimation.py:154-164)
    st[CG_RSNEW] = s;
    st[CG_ITERS] += 1.0;
    if (sigpy ? sqrt(fabs(s)) <= tol : sqrt(fabs(s)) < tol) {
      st[CG_DONE] = 1.0;
    } else {
      st[CG_BETA] = s / st[CG_RSOLD];
      st[CG_RSOLD] = s;
    }
  }
}

// alpha = rsold / pAp; x += alpha p; r -= alpha Ap; partial sums of P |r|^2
__global__ void __launch_bounds__(kVecThreads) k_cg_update(float2* __restrict__ x, float2* __restrict__ r,
                                                           const float2* __restrict__ p, const float2* __restrict__ Ap,
                                                           const float* __restrict__ P, long long n,
                                                           double* __restrict__ st, double* __restrict__ partial) {
  __shared__ double sh[kVecThreads / 64];
  if (st[CG_DONE] != 0.0) return;
  const double ad = st[CG_RSOLD] / st[CG_PAP];
  const float alpha = (float)ad;
  if (blockIdx.x == 0 && threadIdx.x == 0) st[CG_ALPHA] = ad;
  double v = 0.0;
  for (long long i = (long long)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kVecThreads) {
    const float2 pi = p[i];
    const float2 api = Ap[i];
    float2 xi = x[i];
    xi.x = fmaf(alpha, pi.x, xi.x);
    xi.y = fmaf(alpha, pi.y, xi.y);
    x[i] = xi;
    float2 ri = r[i];
    ri.x = fmaf(-alpha, api.x, ri.x);
    ri.y = fmaf(-alpha, api.y, ri.y);
    r[i] = ri;
    const float w = P ? P[i] : 1.0f;
    v += (double)w * ((double)ri.x * ri.x + (double)ri.y * ri.y);
  }
  const double s = wg_sum<kVecThreads>(v, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// y += alpha * yp with the same alpha as the x update (skipped once done):
// keeps E T x up to date as x moves along p, from E T p of the normal operator.
__global__ void __launch_bounds__(kVecThreads) k_cg_track(float2* __restrict__ y, const float2* __restrict__ yp,
                                                          long long n, const double* __restrict__ st) {
  if (st[CG_DONE] != 0.0) return;
  const float alpha = (float)(st[CG_RSOLD] / st[CG_PAP]);
  for (long long i = (long long)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kVecThreads) {
    const float2 a = yp[i];
    float2 b = y[i];
    b.x = fmaf(alpha, a.x, b.x);
    b.y = fmaf(alpha, a.y, b.y);
    y[i] = b;
  }
}

// out = a - b
__global__ void k_csub(float2* __restrict__ out, const float2* __restrict__ a, const float2* __restrict__ b,
                       long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = c_sub(a[i], b[i]);
}

// p = P r + beta p
__global__ void __launch_bounds__(kVecThreads) k_cg_pupdate(float2* __restrict__ p, const float2* __restrict__ r,
                                                            const float* __restrict__ P, long long n,
                                                            const double* __restrict__ st) {
  if (st[CG_DONE] != 0.0) return;
  const float beta = (float)st[CG_BETA];
  for (long long i = (long long)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kVecThreads) {
    const float w = P ? P[i] : 1.0f;
    const float2 ri = r[i];
    float2 pi = p[i];
    pi.x = fmaf(beta, pi.x, w * ri.x);
    pi.y = fmaf(beta, pi.y, w * ri.y);
    p[i] = pi;
  }
}

__global__ void k_mask_mul(float2* __restrict__ x, const unsigned char* __restrict__ m, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    if (!m[i]) x[i] = make_float2(0.f, 0.f);
}

// ---------------------------------------------------------------- LM
constexpr int kStats = 28;  // e0, g[6], H upper triangle (21)

// Per shot: sums over its gathered samples of |r|^2, Re(conj(r) J_p),
// Re(conj(J_i) J_j) (i <= j).  J columns at J + p*jstride.
__global__ void __launch_bounds__(kVecThreads) k_lm_gram(const float2* __restrict__ r, const float2* __restrict__ J,
                                                         long long jstride, const int* __restrict__ off,
                                                         const int* __restrict__ cnt, int per_sample,
                                                         double* __restrict__ partial) {
  __shared__ double sh[kStats][kVecThreads / 64];
  const int b = blockIdx.y;
  const long long base = (long long)off[b] * per_sample;
  const long long n = (long long)cnt[b] * per_sample;
  double acc[kStats];
#pragma unroll
  for (int k = 0; k < kStats; ++k) acc[k] = 0.0;
  for (long long i = (long long)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kVecThreads) {
    const float2 ri = r[base + i];
    float2 jv[6];
#pragma unroll
    for (int p = 0; p < 6; ++p) jv[p] = J[p * jstride + base + i];
    acc[0] += (double)ri.x * ri.x + (double)ri.y * ri.y;
#pragma unroll
    for (int p = 0; p < 6; ++p) acc[1 + p] += (double)ri.x * jv[p].x + (double)ri.y * jv[p].y;
    int k = 7;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int c = a; c < 6; ++c) acc[k++] += (double)jv[a].x * jv[c].x + (double)jv[a].y * jv[c].y;
  }
#pragma unroll
  for (int k = 0; k < kStats; ++k) {
    double v = wave_sum(acc[k]);
    if ((threadIdx.x & 63) == 0) sh[k][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < kStats) {
    double s = 0.0;
    for (int w = 0; w < kVecThreads / 64; ++w) s += sh[threadIdx.x][w];
    partial[((long long)b * gridDim.x + blockIdx.x) * kStats + threadIdx.x] = s;
  }
}

__global__ void k_lm_gram_finalize(const double* __restrict__ partial, int nblk, double* __restrict__ stats) {
  const int b = blockIdx.x;
  const int k = threadIdx.x;
  if (k >= kStats) return;
  double s = 0.0;
  for (int i = 0; i < nblk; ++i) s += partial[((long long)b * nblk + i) * kStats + k];
  stats[(long long)b * kStats + k] = s;
}

// H delta = g with H = Gram + damp*I (transform_estimation.py:60-71), then
// candidate = theta - delta (transform_estimation.py:125).  Singular systems
// produce non-finite deltas, as a failed LU would.
// d = (Gram + damp I)^-1 rhs for the Gram matrix packed in a stats row
// (Gaussian elimination with partial pivoting, FP64).
__device__ void solve6(const double* __restrict__ st, double damp, const double* rhs, double* d) {
  double H[6][7];
  int k = 7;
  for (int a = 0; a < 6; ++a)
    for (int c = a; c < 6; ++c) {
      H[a][c] = st[k];
      H[c][a] = st[k];
      ++k;
    }
  for (int a = 0; a < 6; ++a) {
    H[a][a] += damp;
    H[a][6] = rhs[a];
  }
  for (int col = 0; col < 6; ++col) {
    int piv = col;
    double best = fabs(H[col][col]);
    for (int row = col + 1; row < 6; ++row)
      if (fabs(H[row][col]) > best) {
        best = fabs(H[row][col]);
        piv = row;
      }
    if (piv != col)
      for (int j = 0; j < 7; ++j) {
        const double tmp = H[col][j];
        H[col][j] = H[piv][j];
        H[piv][j] = tmp;
      }
    const double inv = 1.0 / H[col][col];
    for (int row = col + 1; row < 6; ++row) {
      const double f = H[row][col] * inv;
      for (int j = col; j < 7; ++j) H[row][j] -= f * H[col][j];
    }
  }
  for (int row = 5; row >= 0; --row) {
    double s = H[row][6];
    for (int j = row + 1; j < 6; ++j) s -= H[row][j] * d[j];
    d[row] = s / H[row][row];
  }
}

__global__ void k_lm_solve(const double* __restrict__ stats, const double* __restrict__ damp,
                           const double* __restrict__ theta, int th_stride, double* __restrict__ cand,
                           double* __restrict__ delta_out, int nbatch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbatch) return;
  const double* st = stats + (long long)b * kStats;
  double d[6];
  solve6(st, damp[b], st + 1, d);
  for (int a = 0; a < 6; ++a) {
    cand[(long long)b * 6 + a] = theta[(long long)b * th_stride + a] - d[a];
    if (delta_out) delta_out[(long long)b * 6 + a] = d[a];
  }
}

// Geodesic acceleration (transform_estimation.py:78-100), directional
// derivative J'v of the Jacobian along v = -delta, formed in image space:
// out_b (+)= coef * v_b[p] * d_b for one derivative image d_b per shot (the
// column-p image at a perturbed theta), coef = +-1/(2h) for a central
// difference.  Encoding the accumulated image once gives J'v (the encoding
// is linear).  Accumulation is FP32.
__global__ void __launch_bounds__(kVecThreads) k_lm_ga_accumulate(const float2* __restrict__ d, long long d_bs,
                                                                  long long n, const double* __restrict__ delta,
                                                                  int p_idx, double coef, int accumulate,
                                                                  float2* __restrict__ out, long long out_bs) {
  const int b = blockIdx.y;
  const float w = (float)(-coef * delta[(long long)b * 6 + p_idx]);
  const float2* db = d + (long long)b * d_bs;
  float2* ob = out + (long long)b * out_bs;
  for (long long i = (long long)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kVecThreads) {
    const float2 a = db[i];
    float2 o = accumulate ? ob[i] : make_float2(0.f, 0.f);
    o.x = fmaf(w, a.x, o.x);
    o.y = fmaf(w, a.y, o.y);
    ob[i] = o;
  }
}

// transform_estimation.py:102-117: rhs_p = -Re<J'v, J_p> (stats_dir row
// entries 1..6 from the Gram kernel run on J'v), a = (Gram + damp I)^-1 rhs,
// v_ga = v + a/2 when |a| <= |v| (else v), candidate = theta + v_ga.
__global__ void k_lm_ga_solve(const double* __restrict__ stats, const double* __restrict__ stats_dir,
                              const double* __restrict__ damp, const double* __restrict__ theta, int th_stride,
                              const double* __restrict__ delta, double* __restrict__ cand, int nbatch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbatch) return;
  double rhs[6], a[6], v[6];
  for (int p = 0; p < 6; ++p) {
    rhs[p] = -stats_dir[(long long)b * kStats + 1 + p];
    v[p] = -delta[(long long)b * 6 + p];
  }
  solve6(stats + (long long)b * kStats, damp[b], rhs, a);
  double na = 0.0, nv = 0.0;
  for (int p = 0; p < 6; ++p) {
    na += a[p] * a[p];
    nv += v[p] * v[p];
  }
  const bool take = sqrt(na) <= 1.0 * sqrt(nv);
  for (int p = 0; p < 6; ++p)
    cand[(long long)b * 6 + p] = theta[(long long)b * th_stride + p] + v[p] + (take ? 0.5 * a[p] : 0.0);
}

// transform_estimation.py:138-144
__global__ void k_lm_accept(const double* __restrict__ stats, const double* __restrict__ e1, double* __restrict__ theta,
                            int th_stride, const double* __restrict__ cand, double* __restrict__ damp,
                            unsigned char* __restrict__ flags, unsigned char* __restrict__ accepted, double tol,
                            int nbatch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbatch) return;
  const double e0 = stats[(long long)b * kStats];
  const bool acc = e1[b] < e0;
  if (acc) {
    bool conv = true;
    for (int a = 0; a < 6; ++a) conv = conv && fabs(theta[(long long)b * th_stride + a] - cand[b * 6 + a]) < tol;
    if (conv && flags) flags[b] = 1;
    for (int a = 0; a < 6; ++a) theta[(long long)b * th_stride + a] = cand[b * 6 + a];
    damp[b] = fmax(damp[b] / 5.0, 1e-4);
  } else {
    damp[b] = fmin(damp[b] * 1.5, 1e16);
  }
  if (accepted) accepted[b] = acc ? 1 : 0;
}

// mean over shots (fixed order) and subtraction (transform_estimation.py:147-150)
__global__ void k_lm_mean(double* __restrict__ theta, int th_stride, int nshots, double* __restrict__ mean) {
  __shared__ double m[6];
  if (threadIdx.x < 6) {
    double s = 0.0;
    for (int b = 0; b < nshots; ++b) s += theta[(long long)b * th_stride + threadIdx.x];
    m[threadIdx.x] = s / nshots;
    mean[threadIdx.x] = m[threadIdx.x];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nshots * 6; i += blockDim.x) theta[(long long)(i / 6) * th_stride + (i % 6)] -= m[i % 6];
}

}  // namespace me

using namespace me;

extern "C" {

int me_vec_partials_size(void) { return kVecBlocks; }

int me_dot_re(const void* a, const void* b, const float* w, long long n, double* partials, double* out,
              hipStream_t stream) {
  ME_REQUIRE(a && b && partials && out && n >= 0, "dot_re: bad argument");
  hipLaunchKernelGGL(k_dot_re, dim3(kVecBlocks), dim3(kVecThreads), 0, stream, (const float2*)a, (const float2*)b, w,
                     n, partials);
  hipLaunchKernelGGL(k_finalize_sum, dim3(1), dim3(kVecThreads), 0, stream, partials, kVecBlocks, out);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_cg_init(const void* b, const void* Ap, void* r, void* p, const float* P, long long n, double* state,
               double* partials, hipStream_t stream) {
  ME_REQUIRE(b && Ap && r && p && state && partials, "cg_init: null argument");
  hipLaunchKernelGGL(k_cg_init, dim3(kVecBlocks), dim3(kVecThreads), 0, stream, (const float2*)b, (const float2*)Ap,
                     (float2*)r, (float2*)p, P, n, partials);
  hipLaunchKernelGGL(k_cg_finalize, dim3(1), dim3(kVecThreads), 0, stream, partials, kVecBlocks, state, 0, 0.0);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_cg_step(void* x, void* r, void* p, const void* Ap, const float* P, long long n, double tol, double* state,
               double* partials, void* yx, const void* yp, long long ny, hipStream_t stream) {
  ME_REQUIRE(x && r && p && Ap && state && partials, "cg_step: null argument");
  ME_REQUIRE(!yx == !yp, "cg_step: tracking needs both yx and yp");
  hipLaunchKernelGGL(k_dot_re, dim3(kVecBlocks), dim3(kVecThreads), 0, stream, (const float2*)p, (const float2*)Ap,
                     (const float*)nullptr, n, partials);
  hipLaunchKernelGGL(k_cg_finalize, dim3(1), dim3(kVecThreads), 0, stream, partials, kVecBlocks, state, 1, 0.0);
  hipLaunchKernelGGL(k_cg_update, dim3(kVecBlocks), dim3(kVecThreads), 0, stream, (float2*)x, (float2*)r,
                     (const float2*)p, (const float2*)Ap, P, n, state, partials);
  if (yx)
    hipLaunchKernelGGL(k_cg_track, dim3(kVecBlocks), dim3(kVecThreads), 0, stream, (float2*)yx, (const float2*)yp, ny,
                       (const double*)state);
  hipLaunchKernelGGL(k_cg_finalize, dim3(1), dim3(kVecThreads), 0, stream, partials, kVecBlocks, state, 2, tol);
  hipLaunchKernelGGL(k_cg_pupdate, dim3(kVecBlocks), dim3(kVecThreads), 0, stream, (float2*)p, (const float2*)r, P, n,
                     (const double*)state);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

long long me_dot_batched_partials_size(int nbatch) { return (long long)kDotBatchBlocks * (nbatch > 0 ? nbatch : 0); }

int me_dot_re_batched(const void* a, long long a_bstride, const void* c, long long c_bstride, long long n, int nbatch,
                      double* partials, double* out, long long out_stride, hipStream_t stream) {
  ME_REQUIRE(a && c && partials && out && n >= 0, "dot_re_batched: bad argument");
  ME_REQUIRE(nbatch >= 1 && nbatch <= 65535, "dot_re_batched: bad batch count");
  hipLaunchKernelGGL(k_dot_re_batched, dim3(kDotBatchBlocks, nbatch), dim3(kVecThreads), 0, stream, (const float2*)a,
                     a_bstride, (const float2*)c, c_bstride, n, partials);
  hipLaunchKernelGGL(k_finalize_batched, dim3(nbatch), dim3(kVecThreads), 0, stream, partials, kDotBatchBlocks, out,
                     out_stride);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_sum_batch(const void* x, long long x_bstride, int nbatch, long long n, void* out, int accumulate,
                 hipStream_t stream) {
  ME_REQUIRE(x && out && nbatch >= 1 && n >= 0, "sum_batch: bad argument");
  hipLaunchKernelGGL(k_sum_batch, dim3(kVecBlocks), dim3(kVecThreads), 0, stream, (const float2*)x, x_bstride, nbatch,
                     n, (float2*)out, accumulate);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_axpy_real(void* y, double a, const void* x, long long n, hipStream_t stream) {
  ME_REQUIRE(y && x && n >= 0, "axpy_real: bad argument");
  hipLaunchKernelGGL(k_axpy_real, dim3(kVecBlocks), dim3(kVecThreads), 0, stream, (float2*)y, (float)a,
                     (const float2*)x, n);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_csub(void* out, const void* a, const void* b, long long n, hipStream_t stream) {
  ME_REQUIRE(out && a && b && n >= 0, "csub: bad argument");
  hipLaunchKernelGGL(k_csub, dim3(kVecBlocks), dim3(kVecThreads), 0, stream, (float2*)out, (const float2*)a,
                     (const float2*)b, n);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_mask_mul(void* x, const unsigned char* mask, long long n, hipStream_t stream) {
  ME_REQUIRE(x && mask, "mask_mul: null argument");
  hipLaunchKernelGGL(k_mask_mul, dim3(kVecBlocks), dim3(kVecThreads), 0, stream, (float2*)x, mask, n);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

long long me_lm_partials_size(int nbatch) { return (long long)nbatch * 64 * kStats; }

int me_lm_gram(const void* r, const void* J, long long jstride, const int* shot_off, const int* shot_cnt,
               int per_sample, int nbatch, double* partials, double* stats, hipStream_t stream) {
  ME_REQUIRE(r && J && shot_off && shot_cnt && partials && stats && nbatch >= 1, "lm_gram: bad argument");
  const int nblk = 64;
  hipLaunchKernelGGL(k_lm_gram, dim3(nblk, nbatch), dim3(kVecThreads), 0, stream, (const float2*)r, (const float2*)J,
                     jstride, shot_off, shot_cnt, per_sample, partials);
  hipLaunchKernelGGL(k_lm_gram_finalize, dim3(nbatch), dim3(64), 0, stream, partials, nblk, stats);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_lm_solve(const double* stats, const double* damp, const double* theta, int theta_stride, double* cand,
                double* delta, int nbatch, hipStream_t stream) {
  ME_REQUIRE(stats && damp && theta && cand && nbatch >= 1, "lm_solve: bad argument");
  hipLaunchKernelGGL(k_lm_solve, dim3((nbatch + 63) / 64), dim3(64), 0, stream, stats, damp, theta, theta_stride, cand,
                     delta, nbatch);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_lm_ga_accumulate(const void* d, long long d_bstride, long long n, int nbatch, const double* delta, int p_idx,
                        double coef, int accumulate, void* out, long long out_bstride, hipStream_t stream) {
  ME_REQUIRE(d && delta && out && n >= 1, "lm_ga_accumulate: bad argument");
  ME_REQUIRE(p_idx >= 0 && p_idx < 6, "lm_ga_accumulate: p_idx out of range");
  ME_REQUIRE(nbatch >= 1 && nbatch <= 65535, "lm_ga_accumulate: bad batch count");
  const long long blocks = std::min<long long>((n + kVecThreads - 1) / kVecThreads, 1024);
  hipLaunchKernelGGL(k_lm_ga_accumulate, dim3((unsigned)blocks, nbatch), dim3(kVecThreads), 0, stream,
                     (const float2*)d, d_bstride, n, delta, p_idx, coef, accumulate ? 1 : 0, (float2*)out,
                     out_bstride);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_lm_ga_solve(const double* stats, const double* stats_dir, const double* damp, const double* theta,
                   int theta_stride, const double* delta, double* cand, int nbatch, hipStream_t stream) {
  ME_REQUIRE(stats && stats_dir && damp && theta && delta && cand && nbatch >= 1, "lm_ga_solve: bad argument");
  hipLaunchKernelGGL(k_lm_ga_solve, dim3((nbatch + 63) / 64), dim3(64), 0, stream, stats, stats_dir, damp, theta,
                     theta_stride, delta, cand, nbatch);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_lm_accept(const double* stats, const double* e1, double* theta, int theta_stride, const double* cand,
                 double* damp, unsigned char* flags, unsigned char* accepted, double tol, int nbatch,
                 hipStream_t stream) {
  ME_REQUIRE(stats && e1 && theta && cand && damp && nbatch >= 1, "lm_accept: bad argument");
  hipLaunchKernelGGL(k_lm_accept, dim3((nbatch + 63) / 64), dim3(64), 0, stream, stats, e1, theta, theta_stride, cand,
                     damp, flags, accepted, tol, nbatch);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_lm_mean(double* theta, int theta_stride, int nshots, double* mean, hipStream_t stream) {
  ME_REQUIRE(theta && mean && nshots >= 1, "lm_mean: bad argument");
  hipLaunchKernelGGL(k_lm_mean, dim3(1), dim3(256), 0, stream, theta, theta_stride, nshots, mean);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
