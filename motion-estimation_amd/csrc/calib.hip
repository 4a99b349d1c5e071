// Calibration preprocessing before the hot path (noise_whiten.py:5-14,
// coil_compression.py:6-46): coil Gram matrices, the whitening matrix from
// a Cholesky factor, a Hermitian eigendecomposition (for the SVD of the
// Hermitian coil-compression matrix) and coil mixing.  Small C x C algebra
// runs in FP64 inside one workgroup with the matrices in LDS; the Gram sums
// over voxels use FP64 per-workgroup partials reduced in a fixed order.
#include "me_device.h"
#include "me_internal.h"
#include "../../include/motion_estimation_amd.h"

namespace me {

constexpr int kMaxCalibCoils = 64;
constexpr int kGramThreads = 256;
constexpr int kGramTile = 64;     // voxels staged per step
constexpr int kGramBlocks = 512;  // workgroups of the Gram sum
constexpr int kGramPairsPerThread = (kMaxCalibCoils * (kMaxCalibCoils + 1) / 2 + kGramThreads - 1) / kGramThreads;

struct Z {  // FP64 complex
  double re, im;
};
__device__ __forceinline__ Z zmul(Z a, Z b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Z zconj(Z a) { return {a.re, -a.im}; }
__device__ __forceinline__ Z zsub(Z a, Z b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ Z zadd(Z a, Z b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Z zscale(Z a, double s) { return {a.re * s, a.im * s}; }

// Partial sums of G_ij = sum_v w(v) x_i(v) conj(x_j(v)) for i <= j over this
// workgroup's voxels; partial[(blk * npairs + pair) * 2 + {0, 1}].
__global__ void __launch_bounds__(kGramThreads) k_coil_gram(const float2* __restrict__ x, int C, long long n,
                                                           const float* __restrict__ w, double* __restrict__ partial) {
  __shared__ float2 tile[kGramTile * kMaxCalibCoils];
  __shared__ float wt[kGramTile];
  const int npairs = C * (C + 1) / 2;
  int pi[kGramPairsPerThread], pj[kGramPairsPerThread];
  double are[kGramPairsPerThread], aim[kGramPairsPerThread];
#pragma unroll
  for (int k = 0; k < kGramPairsPerThread; ++k) {
    int p = threadIdx.x + k * kGramThreads;
    int i = 0, j = 0;
    if (p < npairs) {  // pair index -> (i, j), i <= j, row-major upper triangle
      int rem = p;
      while (rem >= C - i) {
        rem -= C - i;
        ++i;
      }
      j = i + rem;
    }
    pi[k] = i;
    pj[k] = j;
    are[k] = 0.0;
    aim[k] = 0.0;
  }
  const long long ntiles = (n + kGramTile - 1) / kGramTile;
  for (long long tl = blockIdx.x; tl < ntiles; tl += gridDim.x) {
    const long long v0 = tl * kGramTile;
    __syncthreads();
    for (int e = threadIdx.x; e < kGramTile * C; e += kGramThreads) {
      const int c = e / kGramTile, v = e % kGramTile;
      const long long gv = v0 + v;
      tile[v * C + c] = gv < n ? x[(long long)c * n + gv] : make_float2(0.f, 0.f);
    }
    for (int v = threadIdx.x; v < kGramTile; v += kGramThreads) {
      const long long gv = v0 + v;
      wt[v] = gv < n ? (w ? w[gv] : 1.f) : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kGramPairsPerThread; ++k) {
      if (threadIdx.x + k * kGramThreads >= npairs) continue;
      double sr = 0.0, si = 0.0;
      for (int v = 0; v < kGramTile; ++v) {
        const float2 a = tile[v * C + pi[k]], b = tile[v * C + pj[k]];
        const double ww = (double)wt[v];
        // a * conj(b)
        sr += ww * ((double)a.x * b.x + (double)a.y * b.y);
        si += ww * ((double)a.y * b.x - (double)a.x * b.y);
      }
      are[k] += sr;
      aim[k] += si;
    }
  }
#pragma unroll
  for (int k = 0; k < kGramPairsPerThread; ++k) {
    const int p = threadIdx.x + k * kGramThreads;
    if (p < npairs) {
      partial[((long long)blockIdx.x * npairs + p) * 2] = are[k];
      partial[((long long)blockIdx.x * npairs + p) * 2 + 1] = aim[k];
    }
  }
}

// G (C x C, interleaved FP64 complex, full Hermitian) = scale * sum over blocks
__global__ void k_coil_gram_finalize(const double* __restrict__ partial, int nblk, int C, double scale,
                                     double* __restrict__ G) {
  const int npairs = C * (C + 1) / 2;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += gridDim.x * blockDim.x) {
    int i = 0, rem = p;
    while (rem >= C - i) {
      rem -= C - i;
      ++i;
    }
    const int j = i + rem;
    double sr = 0.0, si = 0.0;
    for (int b = 0; b < nblk; ++b) {
      sr += partial[((long long)b * npairs + p) * 2];
      si += partial[((long long)b * npairs + p) * 2 + 1];
    }
    sr *= scale;
    si *= scale;
    G[(i * C + j) * 2] = sr;
    G[(i * C + j) * 2 + 1] = si;
    G[(j * C + i) * 2] = sr;
    G[(j * C + i) * 2 + 1] = -si;
  }
}

// W = (L^-1)^H with L L^H = G + reg I (lower Cholesky factor), as
// noise_whiten.py:8-10 computes it.  One workgroup; matrices in LDS.
__global__ void __launch_bounds__(64) k_whitening(const double* __restrict__ G, int C, double reg,
                                                  double* __restrict__ W) {
  extern __shared__ double lds[];
  Z* L = (Z*)lds;             // C x C
  Z* Li = (Z*)lds + C * C;    // C x C, inverse of L
  const int tid = threadIdx.x;
  for (int e = tid; e < C * C; e += blockDim.x) {
    const int i = e / C, j = e % C;
    Z g{G[e * 2], G[e * 2 + 1]};
    if (i == j) g.re += reg;
    L[e] = g;
    Li[e] = Z{i == j ? 1.0 : 0.0, 0.0};
  }
  __syncthreads();
  // Cholesky, column by column (lower triangle of L overwrites G's)
  for (int j = 0; j < C; ++j) {
    if (tid == 0) {
      double d = L[j * C + j].re;
      for (int k = 0; k < j; ++k) d -= L[j * C + k].re * L[j * C + k].re + L[j * C + k].im * L[j * C + k].im;
      L[j * C + j] = Z{sqrt(d), 0.0};
    }
    __syncthreads();
    const double inv = 1.0 / L[j * C + j].re;
    for (int i = j + 1 + tid; i < C; i += blockDim.x) {
      Z s = L[i * C + j];
      for (int k = 0; k < j; ++k) s = zsub(s, zmul(L[i * C + k], zconj(L[j * C + k])));
      L[i * C + j] = zscale(s, inv);
    }
    __syncthreads();
  }
  // L X = I by forward substitution, one column of X per thread
  for (int col = tid; col < C; col += blockDim.x) {
    for (int i = 0; i < C; ++i) {
      Z s{i == col ? 1.0 : 0.0, 0.0};
      for (int k = col; k < i; ++k) s = zsub(s, zmul(L[i * C + k], Li[k * C + col]));
      Li[i * C + col] = i < col ? Z{0.0, 0.0} : zscale(s, 1.0 / L[i * C + i].re);
    }
  }
  __syncthreads();
  for (int e = tid; e < C * C; e += blockDim.x) {  // W = Li^H
    const int i = e / C, j = e % C;
    const Z v = Li[j * C + i];
    W[e * 2] = v.re;
    W[e * 2 + 1] = -v.im;
  }
}

// Hermitian eigendecomposition A = V diag(lam) V^H by cyclic complex Jacobi
// rotations (FP64, one workgroup, A and V in LDS), eigenvalues descending.
// A rotation of the pair (p, q) with a_pq = |a_pq| e^{i phi} is J = D R,
// D = diag(1, e^{-i phi}) making the 2x2 block real, R the real Jacobi
// rotation with t = sign(tau) / (|tau| + sqrt(1 + tau^2)),
// tau = (a_qq - a_pp) / (2 |a_pq|).
__global__ void __launch_bounds__(64) k_herm_eig(const double* __restrict__ Ain, int C, int max_sweeps,
                                                 double* __restrict__ lam, double* __restrict__ Vout) {
  extern __shared__ double lds[];
  Z* A = (Z*)lds;
  Z* V = (Z*)lds + C * C;
  __shared__ double red[64];
  __shared__ Z rot[4];
  __shared__ int order[kMaxCalibCoils];
  const int tid = threadIdx.x;
  for (int e = tid; e < C * C; e += blockDim.x) {
    A[e] = Z{Ain[e * 2], Ain[e * 2 + 1]};
    V[e] = Z{(e / C) == (e % C) ? 1.0 : 0.0, 0.0};
  }
  __syncthreads();
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int e = tid; e < C * C; e += blockDim.x) {
      const double m = A[e].re * A[e].re + A[e].im * A[e].im;
      tot += m;
      if (e / C != e % C) off += m;
    }
    red[tid] = off;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int k = 0; k < (int)blockDim.x; ++k) s += red[k];
      red[0] = s;
    }
    __syncthreads();
    const double off_all = red[0];
    __syncthreads();
    red[tid] = tot;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int k = 0; k < (int)blockDim.x; ++k) s += red[k];
      red[0] = s;
    }
    __syncthreads();
    const double tot_all = red[0];
    __syncthreads();
    if (off_all <= 1e-28 * tot_all) break;
    for (int p = 0; p < C - 1; ++p)
      for (int q = p + 1; q < C; ++q) {
        if (tid == 0) {
          const Z c = A[p * C + q];
          const double ac = sqrt(c.re * c.re + c.im * c.im);
          if (ac < 1e-300) {
            rot[0] = Z{1.0, 0.0};
            rot[1] = Z{0.0, 0.0};
            rot[2] = Z{0.0, 0.0};
            rot[3] = Z{1.0, 0.0};
          } else {
            const double a = A[p * C + p].re, b = A[q * C + q].re;
            const double tau = (b - a) / (2.0 * ac);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            const double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
            const Z ephc{c.re / ac, -c.im / ac};  // e^{-i phi}
            rot[0] = Z{cs, 0.0};                  // J_pp
            rot[1] = Z{sn, 0.0};                  // J_pq
            rot[2] = zscale(ephc, -sn);           // J_qp
            rot[3] = zscale(ephc, cs);            // J_qq
          }
        }
        __syncthreads();
        const Z jpp = rot[0], jpq = rot[1], jqp = rot[2], jqq = rot[3];
        for (int k = tid; k < C; k += blockDim.x) {  // columns p, q of A and V
          const Z xp = A[k * C + p], xq = A[k * C + q];
          A[k * C + p] = zadd(zmul(xp, jpp), zmul(xq, jqp));
          A[k * C + q] = zadd(zmul(xp, jpq), zmul(xq, jqq));
          const Z vp = V[k * C + p], vq = V[k * C + q];
          V[k * C + p] = zadd(zmul(vp, jpp), zmul(vq, jqp));
          V[k * C + q] = zadd(zmul(vp, jpq), zmul(vq, jqq));
        }
        __syncthreads();
        for (int k = tid; k < C; k += blockDim.x) {  // rows p, q of A
          const Z yp = A[p * C + k], yq = A[q * C + k];
          A[p * C + k] = zadd(zmul(zconj(jpp), yp), zmul(zconj(jqp), yq));
          A[q * C + k] = zadd(zmul(zconj(jpq), yp), zmul(zconj(jqq), yq));
        }
        __syncthreads();
      }
  }
  if (tid == 0) {  // descending order (stable selection sort)
    for (int k = 0; k < C; ++k) order[k] = k;
    for (int a = 0; a < C; ++a) {
      int best = a;
      for (int b = a + 1; b < C; ++b)
        if (A[order[b] * C + order[b]].re > A[order[best] * C + order[best]].re) best = b;
      const int tmp = order[a];
      order[a] = order[best];
      order[best] = tmp;
    }
  }
  __syncthreads();
  for (int k = tid; k < C; k += blockDim.x) lam[k] = A[order[k] * C + order[k]].re;
  for (int e = tid; e < C * C; e += blockDim.x) {
    const int i = e / C, k = e % C;  // row i, output column k
    const Z v = V[i * C + order[k]];
    Vout[e * 2] = v.re;
    Vout[e * 2 + 1] = v.im;
  }
}

// y[o][v] = sum_c M[o][c] x[c][v]  (M: cout x cin complex, row-major).
// Outputs in register blocks of 8; each block re-reads the voxel's inputs
// (cache hits after the first block).
// y[o](v) = sum_c M[o][c] x[c](v): each thread reads its voxel's cin inputs
// once into registers (MAXC >= cin), then forms the outputs 8 at a time from
// the LDS copy of M, so x is read once and y written once (2 C V bytes).
template <int MAXC>
__global__ void __launch_bounds__(256) k_coil_mix(const float2* __restrict__ M, int cout, int cin,
                                                  const float2* __restrict__ x, long long n, float2* __restrict__ y) {
  __shared__ float2 m[kMaxCalibCoils * kMaxCalibCoils];
  for (int e = threadIdx.x; e < cout * cin; e += blockDim.x) m[e] = M[e];
  __syncthreads();
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x) {
    float2 xin[MAXC];
#pragma unroll
    for (int c = 0; c < MAXC; ++c) xin[c] = c < cin ? x[(long long)c * n + v] : make_float2(0.f, 0.f);
    for (int o0 = 0; o0 < cout; o0 += 8) {
      float2 acc[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[j] = make_float2(0.f, 0.f);
#pragma unroll
      for (int c = 0; c < MAXC; ++c) {
        if (c < cin) {
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const int o = o0 + j < cout ? o0 + j : cout - 1;
            acc[j] = c_fma(m[o * cin + c], xin[c], acc[j]);
          }
        }
      }
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (o0 + j < cout) y[(long long)(o0 + j) * n + v] = acc[j];
    }
  }
}

}  // namespace me

using namespace me;

extern "C" {

long long me_coil_gram_partials_size(int ncoils) {
  if (ncoils < 1 || ncoils > kMaxCalibCoils) return 0;
  return (long long)kGramBlocks * ncoils * (ncoils + 1);  // 2 doubles per pair
}

int me_coil_gram(const void* x, int ncoils, long long n, const float* w, double scale, double* partials,
                 double* gram, hipStream_t stream) {
  ME_REQUIRE(x && partials && gram && n >= 1, "coil_gram: bad argument");
  ME_REQUIRE(ncoils >= 1 && ncoils <= kMaxCalibCoils, "coil_gram: 1 to 64 coils");
  hipLaunchKernelGGL(k_coil_gram, dim3(kGramBlocks), dim3(kGramThreads), 0, stream, (const float2*)x, ncoils, n, w,
                     partials);
  hipLaunchKernelGGL(k_coil_gram_finalize, dim3(8), dim3(256), 0, stream, partials, kGramBlocks, ncoils, scale, gram);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_whitening_matrix(const double* gram, int ncoils, double reg, double* W, hipStream_t stream) {
  ME_REQUIRE(gram && W, "whitening_matrix: null argument");
  ME_REQUIRE(ncoils >= 1 && ncoils <= kMaxCalibCoils, "whitening_matrix: 1 to 64 coils");
  const size_t shmem = (size_t)2 * ncoils * ncoils * 2 * sizeof(double);
  ME_TRY_HIP(hipFuncSetAttribute((const void*)k_whitening, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  hipLaunchKernelGGL(k_whitening, dim3(1), dim3(64), shmem, stream, gram, ncoils, reg, W);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_hermitian_eig(const double* A, int n, int max_sweeps, double* eigvals, double* eigvecs, hipStream_t stream) {
  ME_REQUIRE(A && eigvals && eigvecs, "hermitian_eig: null argument");
  ME_REQUIRE(n >= 1 && n <= kMaxCalibCoils, "hermitian_eig: order 1 to 64");
  const size_t shmem = (size_t)2 * n * n * 2 * sizeof(double);
  ME_TRY_HIP(hipFuncSetAttribute((const void*)k_herm_eig, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  hipLaunchKernelGGL(k_herm_eig, dim3(1), dim3(64), shmem, stream, A, n, max_sweeps, eigvals, eigvecs);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_coil_mix(const void* M, int cout, int cin, const void* x, long long n, void* y, hipStream_t stream) {
  ME_REQUIRE(M && x && y && n >= 0, "coil_mix: bad argument");
  ME_REQUIRE(cin >= 1 && cin <= kMaxCalibCoils && cout >= 1 && cout <= kMaxCalibCoils, "coil_mix: 1 to 64 coils");
  ME_REQUIRE(x != y, "coil_mix: output must not alias the input");
  const long long blocks = (n + 255) / 256;
  const dim3 grid((unsigned)(blocks < 8192 ? (blocks > 0 ? blocks : 1) : 8192));
  if (cin <= 16)
    hipLaunchKernelGGL(k_coil_mix<16>, grid, dim3(256), 0, stream, (const float2*)M, cout, cin, (const float2*)x, n,
                       (float2*)y);
  else if (cin <= 32)
    hipLaunchKernelGGL(k_coil_mix<32>, grid, dim3(256), 0, stream, (const float2*)M, cout, cin, (const float2*)x, n,
                       (float2*)y);
  else
    hipLaunchKernelGGL(k_coil_mix<64>, grid, dim3(256), 0, stream, (const float2*)M, cout, cin, (const float2*)x, n,
                       (float2*)y);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
