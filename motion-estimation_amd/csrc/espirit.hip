// ESPIRiT receive-sensitivity estimation (espirit.py:47-239) on the GPU.
//
//  1. Calibration matrix H (m = K*NRx rows, one column per kx*ky*kz block of
//     the calibration k-space), row k + K*c with k the Fortran-order offset
//     inside the block (espirit.py:9-45).
//  2. G = H H^H in FP64 (tiled; FP32 products are exact in FP64).  Its
//     eigenvectors are H's left singular vectors, its eigenvalues the
//     squared singular values (espirit.py:114).
//  3. Eigendecomposition of G by parallel two-sided Jacobi: each round
//     rotates the m/2 disjoint index pairs of a round-robin schedule at once,
//     G <- J^H G J and V <- V J in one kernel (every 2x2 block of G touched
//     by two pairs transforms independently), sweeps until the off-diagonal
//     Frobenius norm is below tol times the diagonal one.
//  4. The kept subspace (singular values above eig_thresh * the largest, or
//     the leading ones by count) enters only through its projector
//     P = V_r V_r^H.  After the centred zero-padded inverse FFTs of
//     espirit.py:146-198 the coil matrix of voxel v is
//        C(v)_{cc'} = (prod N)^-2 sum_d Q_{cc'}[d] exp(i theta(v) . d),
//        Q_{cc'}[d] = sum_{k - k' = d} P[(k,c), (k',c')],
//     theta_a(v) = 2 pi (v_a - N_a//2) / N_a over the (2k-1)^3 difference box
//     (the padding offset only adds a phase common to C's factors, which
//     cancels).  The kernels below evaluate Cs(v) = (prod N)^2 C(v).
//  5. Per voxel: Cs evaluated separably (z for all voxels, y per image row
//     into LDS, x per voxel), its principal eigenvector by power iteration
//     with lane c of a lane group holding row c, the mask
//     sqrt(lambda_max(C)) * prod(N) / sqrt(prod(kernel)) > mask_thresh, and
//     the phase of coil 0 removed (espirit.py:200-239).
#include <atomic>
#include <vector>

#include "me_device.h"
#include "me_internal.h"
#include "../../include/motion_estimation_amd.h"

namespace me {

// ------------------------------------------------------------ 1. Hankel
// calib: (NKx, NKy, NKz, NRx) C order.  H: m x nb row-major.
__global__ void k_hankel(const float2* __restrict__ calib, int NKx, int NKy, int NKz, int NRx, int kx, int ky,
                         int kz, float2* __restrict__ H) {
  const int nbx = NKx - kx + 1, nby = NKy - ky + 1, nbz = NKz - kz + 1;
  const long long nb = (long long)nbx * nby * nbz;
  const int K = kx * ky * kz;
  const long long total = (long long)K * NRx * nb;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e % nb;
    const long long row = e / nb;
    const int c = (int)(row / K), k = (int)(row % K);
    const int ix = k % kx, iy = (k / kx) % ky, iz = k / (kx * ky);
    const int bx = (int)(b % nbx), by = (int)((b / nbx) % nby), bz = (int)(b / ((long long)nbx * nby));
    H[e] = calib[(((long long)(bx + ix) * NKy + (by + iy)) * NKz + (bz + iz)) * NRx + c];
  }
}

// ------------------------------------------------------------ 2. Gram
// G[i][j] = sum_b X[i][b] conj(X[j][b]) (X: m x n complex64 row-major), FP64
// accumulation, one 32 x 32 output tile per workgroup over the upper
// triangle of tiles, mirrored into the lower one.
constexpr int kGT = 32;
constexpr int kGK = 32;
__global__ void __launch_bounds__(256) k_gram_tiled(const float2* __restrict__ X, int m, long long n,
                                                    double* __restrict__ G) {
  __shared__ float2 a[kGK][kGT + 1];   // a[c][r] = X[i0 + r][b0 + c]
  __shared__ float2 bt[kGK][kGT + 1];  // bt[c][r] = X[j0 + r][b0 + c]
  const int ntile = (m + kGT - 1) / kGT;
  int t = blockIdx.x, ti = 0;
  while (t >= ntile - ti) {
    t -= ntile - ti;
    ++ti;
  }
  const int tj = ti + t;
  const int i0 = ti * kGT, j0 = tj * kGT;
  const int tid = threadIdx.x;
  const int li = tid / 8;        // row of the output tile
  const int lj = (tid % 8) * 4;  // 4 consecutive output columns
  double are[4] = {0, 0, 0, 0}, aim[4] = {0, 0, 0, 0};
  for (long long b0 = 0; b0 < n; b0 += kGK) {
    __syncthreads();
    for (int e = tid; e < kGT * kGK; e += 256) {
      const int r = e / kGK, c = e % kGK;
      const long long bb = b0 + c;
      a[c][r] = (bb < n && i0 + r < m) ? X[(long long)(i0 + r) * n + bb] : make_float2(0.f, 0.f);
      bt[c][r] = (bb < n && j0 + r < m) ? X[(long long)(j0 + r) * n + bb] : make_float2(0.f, 0.f);
    }
    __syncthreads();
#pragma unroll 4
    for (int c = 0; c < kGK; ++c) {
      const float2 x = a[c][li];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float2 y = bt[c][lj + q];
        are[q] += (double)x.x * y.x + (double)x.y * y.y;
        aim[q] += (double)x.y * y.x - (double)x.x * y.y;
      }
    }
  }
  const int i = i0 + li;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = j0 + lj + q;
    if (i < m && j < m) {
      G[((long long)i * m + j) * 2] = are[q];
      G[((long long)i * m + j) * 2 + 1] = aim[q];
      G[((long long)j * m + i) * 2] = are[q];
      G[((long long)j * m + i) * 2 + 1] = -aim[q];
    }
  }
}

// ------------------------------------------------------------ 3. Jacobi
// Pair i of round r in the round-robin schedule over mm (even) indices:
// index mm-1 is fixed, the others rotate.  Indices >= m are padding.
__device__ __forceinline__ void rr_pair(int mm, int r, int i, int& p, int& q) {
  const int n1 = mm - 1;
  if (i == 0) {
    p = r;
    q = n1;
  } else {
    p = (r + i) % n1;
    q = (r - i + n1) % n1;
  }
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
}

struct Rot {  // J = [[cs, sn], [-sn e, cs e]], e = (er, ei)
  double cs, sn, er, ei;
};

// Rotation annihilating G_pq of the 2x2 block [[a, c], [conj(c), b]]:
// J = D R with D = diag(1, conj(c)/|c|) making the block real and R the real
// symmetric Jacobi rotation tau = (b - a) / (2|c|),
// t = sign(tau) / (|tau| + sqrt(1 + tau^2)), cs = 1/sqrt(1 + t^2), sn = t cs.
__global__ void k_jac_params(const double* __restrict__ G, int m, int mm, int r, double* __restrict__ rot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= mm / 2) return;
  int p, q;
  rr_pair(mm, r, i, p, q);
  double cs = 1.0, sn = 0.0, er = 1.0, ei = 0.0;
  if (q < m) {
    const double cre = G[((long long)p * m + q) * 2], cim = G[((long long)p * m + q) * 2 + 1];
    const double ac = sqrt(cre * cre + cim * cim);
    const double a = G[((long long)p * m + p) * 2], b = G[((long long)q * m + q) * 2];
    if (ac > 1e-300 && ac > 1e-17 * sqrt(fabs(a * b))) {
      const double tau = (b - a) / (2.0 * ac);
      const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
      cs = 1.0 / sqrt(1.0 + t * t);
      sn = t * cs;
      er = cre / ac;
      ei = -cim / ac;
    }
  }
  rot[i * 4 + 0] = cs;
  rot[i * 4 + 1] = sn;
  rot[i * 4 + 2] = er;
  rot[i * 4 + 3] = ei;
}

__device__ __forceinline__ Rot load_rot(const double* rot, int i) {
  return Rot{rot[i * 4], rot[i * 4 + 1], rot[i * 4 + 2], rot[i * 4 + 3]};
}
__device__ __forceinline__ bool is_identity(const Rot& J) { return J.sn == 0.0 && J.er == 1.0 && J.ei == 0.0; }

// (x_p, x_q) J: x_p' = cs x_p - sn e x_q, x_q' = sn x_p + cs e x_q
__device__ __forceinline__ void rot_cols(const Rot& J, double& pr, double& pi, double& qr, double& qi) {
  const double qer = qr * J.er - qi * J.ei, qei = qr * J.ei + qi * J.er;
  const double npr = J.cs * pr - J.sn * qer, npi = J.cs * pi - J.sn * qei;
  qr = J.sn * pr + J.cs * qer;
  qi = J.sn * pi + J.cs * qei;
  pr = npr;
  pi = npi;
}
// J^H (y_p; y_q): y_p' = cs y_p - sn conj(e) y_q, y_q' = sn y_p + cs conj(e) y_q
__device__ __forceinline__ void rot_rows(const Rot& J, double& pr, double& pi, double& qr, double& qi) {
  const double qer = qr * J.er + qi * J.ei, qei = qi * J.er - qr * J.ei;
  const double npr = J.cs * pr - J.sn * qer, npi = J.cs * pi - J.sn * qei;
  qr = J.sn * pr + J.cs * qer;
  qi = J.sn * pi + J.cs * qei;
  pr = npr;
  pi = npi;
}

// One Jacobi round.  blockIdx.y < npairs: rows (p_a, q_a) of G against the
// column pair b of this thread; blockIdx.y >= npairs (launched only when V
// rows do not fit LDS): row blockIdx.y - npairs of V.  Consecutive b give
// column indices in two contiguous runs, so the row accesses coalesce.
__global__ void __launch_bounds__(256) k_jac_round(double* __restrict__ G, double* __restrict__ V, int m, int mm,
                                                   int r, const double* __restrict__ rot) {
  const int npairs = mm / 2;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= npairs) return;
  int pb, qb;
  rr_pair(mm, r, b, pb, qb);
  const Rot Jb = load_rot(rot, b);
  const bool vqb = qb < m;
  if ((int)blockIdx.y < npairs) {
    const int a = blockIdx.y;
    int pa, qa;
    rr_pair(mm, r, a, pa, qa);
    const Rot Ja = load_rot(rot, a);
    const bool vqa = qa < m;
    if (is_identity(Ja) && is_identity(Jb)) return;
    double* gpp = G + ((long long)pa * m + pb) * 2;
    double* gpq = G + ((long long)pa * m + qb) * 2;
    double* gqp = G + ((long long)qa * m + pb) * 2;
    double* gqq = G + ((long long)qa * m + qb) * 2;
    double app_r = gpp[0], app_i = gpp[1];
    double apq_r = vqb ? gpq[0] : 0.0, apq_i = vqb ? gpq[1] : 0.0;
    double aqp_r = vqa ? gqp[0] : 0.0, aqp_i = vqa ? gqp[1] : 0.0;
    double aqq_r = (vqa && vqb) ? gqq[0] : 0.0, aqq_i = (vqa && vqb) ? gqq[1] : 0.0;
    rot_cols(Jb, app_r, app_i, apq_r, apq_i);  // row p_a
    rot_cols(Jb, aqp_r, aqp_i, aqq_r, aqq_i);  // row q_a
    rot_rows(Ja, app_r, app_i, aqp_r, aqp_i);  // column p_b
    rot_rows(Ja, apq_r, apq_i, aqq_r, aqq_i);  // column q_b
    gpp[0] = app_r;
    gpp[1] = app_i;
    if (vqb) {
      gpq[0] = apq_r;
      gpq[1] = apq_i;
    }
    if (vqa) {
      gqp[0] = aqp_r;
      gqp[1] = aqp_i;
    }
    if (vqa && vqb) {
      gqq[0] = aqq_r;
      gqq[1] = aqq_i;
    }
  } else {
    const int row = blockIdx.y - npairs;
    if (!vqb || is_identity(Jb)) return;
    double* vp = V + ((long long)row * m + pb) * 2;
    double* vq = V + ((long long)row * m + qb) * 2;
    double pr = vp[0], pi = vp[1], qr = vq[0], qi = vq[1];
    rot_cols(Jb, pr, pi, qr, qi);
    vp[0] = pr;
    vp[1] = pi;
    vq[0] = qr;
    vq[1] = qi;
  }
}

// V <- V J_r0 ... J_{r0+nr-1}: one workgroup per row of V, the row held in
// LDS across nr rounds (rotation parameters of round r0 + k at
// rots + k * npairs * 4), so V makes one HBM round trip per nr rounds.
constexpr int kRoundsPerVPass = 32;
constexpr int kMaxLdsRow = 8192;  // complex FP64 entries of a row in LDS
// rows up to which V is updated in LDS (kMaxLdsRow, lowered only to test
// the in-round global update at small sizes)
static std::atomic<int> g_jac_lds_rows{kMaxLdsRow};
__global__ void __launch_bounds__(256) k_jac_apply_v(double* __restrict__ V, int m, int mm, int r0, int nr,
                                                     const double* __restrict__ rots) {
  extern __shared__ double2 vrow[];
  double2* gv = reinterpret_cast<double2*>(V) + (long long)blockIdx.x * m;
  for (int j = threadIdx.x; j < m; j += blockDim.x) vrow[j] = gv[j];
  __syncthreads();
  const int npairs = mm / 2;
  for (int k = 0; k < nr; ++k) {
    const double* rot = rots + (long long)k * npairs * 4;
    for (int b = threadIdx.x; b < npairs; b += blockDim.x) {
      int p, q;
      rr_pair(mm, r0 + k, b, p, q);
      if (q >= m) continue;
      const Rot J = load_rot(rot, b);
      if (is_identity(J)) continue;
      double2 xp = vrow[p], xq = vrow[q];
      rot_cols(J, xp.x, xp.y, xq.x, xq.y);
      vrow[p] = xp;
      vrow[q] = xq;
    }
    __syncthreads();
  }
  for (int j = threadIdx.x; j < m; j += blockDim.x) gv[j] = vrow[j];
}

constexpr int kNormBlocks = 256;

// partial[2*blk] = sum |off-diagonal|^2, partial[2*blk+1] = sum |diagonal|^2
__global__ void __launch_bounds__(256) k_jac_norms(const double* __restrict__ G, int m, double* __restrict__ partial) {
  __shared__ double sh[2][4];
  double off = 0.0, dia = 0.0;
  const long long n = (long long)m * m;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
    const double v = G[e * 2] * G[e * 2] + G[e * 2 + 1] * G[e * 2 + 1];
    if (e / m == e % m)
      dia += v;
    else
      off += v;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    off += __shfl_down(off, o, 64);
    dia += __shfl_down(dia, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = off;
    sh[1][threadIdx.x >> 6] = dia;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partial[blockIdx.x * 2] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    partial[blockIdx.x * 2 + 1] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  }
}

__global__ void k_identity(double* __restrict__ V, int m) {
  const long long n = (long long)m * m;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    V[e * 2] = (e / m == e % m) ? 1.0 : 0.0;
    V[e * 2 + 1] = 0.0;
  }
}

__global__ void k_diag(const double* __restrict__ G, int m, double* __restrict__ lam) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) lam[i] = G[((long long)i * m + i) * 2];
}

// X (m x r complex64) = columns idx[0..r) of V (m rows of ldv FP64 complex)
__global__ void k_gather_columns(const double* __restrict__ V, int m, int ldv, const int* __restrict__ idx, int r,
                                 float2* __restrict__ X) {
  const long long n = (long long)m * r;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const int row = (int)(e / r), j = (int)(e % r);
    const long long s = ((long long)row * ldv + idx[j]) * 2;
    X[e] = make_float2((float)V[s], (float)V[s + 1]);
  }
}

// ------------------------------------------------------------ 4. Q sums
// Q[c][c2][dx][dy][dz] = sum over block offsets i, i' with i - i' = d of
// P[i + K c][i' + K c2], d_a in [-(k_a-1), k_a-1].
__global__ void k_espirit_q(const double* __restrict__ P, int NRx, int kx, int ky, int kz, float2* __restrict__ Q) {
  const int K = kx * ky * kz;
  const int m = K * NRx;
  const int Dx = 2 * kx - 1, Dy = 2 * ky - 1, Dz = 2 * kz - 1;
  const long long nd = (long long)Dx * Dy * Dz;
  const long long total = (long long)NRx * NRx * nd;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long di = e % nd;
    const int c2 = (int)((e / nd) % NRx), c = (int)(e / (nd * NRx));
    const int dz = (int)(di % Dz) - (kz - 1);
    const int dy = (int)((di / Dz) % Dy) - (ky - 1);
    const int dx = (int)(di / ((long long)Dz * Dy)) - (kx - 1);
    double sr = 0.0, si = 0.0;
    for (int iz = max(0, dz); iz < min(kz, kz + dz); ++iz)
      for (int iy = max(0, dy); iy < min(ky, ky + dy); ++iy)
        for (int ix = max(0, dx); ix < min(kx, kx + dx); ++ix) {
          const int k = ix + kx * (iy + ky * iz);
          const int k2 = (ix - dx) + kx * ((iy - dy) + ky * (iz - dz));
          const long long idx = ((long long)(k + K * c) * m + (k2 + K * c2)) * 2;
          sr += P[idx];
          si += P[idx + 1];
        }
    Q[e] = make_float2((float)sr, (float)si);
  }
}

// exp(i 2 pi (v - N//2) d / N), the integer product reduced modulo N first
__device__ __forceinline__ float2 esp_phase(int v, int N, int d) {
  const long long t = ((long long)(v - N / 2) * d) % N;
  return cis2pi(frac_half((double)t / (double)N));
}

// Az[z][c][c2][dx][dy] = sum_dz Q[c][c2][dx][dy][dz] exp(i theta_z(z) dz)
__global__ void k_espirit_az(const float2* __restrict__ Q, int NRx, int kx, int ky, int kz, int Nz,
                             float2* __restrict__ Az) {
  const int Dx = 2 * kx - 1, Dy = 2 * ky - 1, Dz = 2 * kz - 1;
  const long long per = (long long)NRx * NRx * Dx * Dy;
  const long long total = per * Nz;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int z = (int)(e / per);
    const long long rest = e % per;
    float2 acc = make_float2(0.f, 0.f);
    for (int j = 0; j < Dz; ++j) acc = c_fma(Q[rest * Dz + j], esp_phase(z, Nz, j - (kz - 1)), acc);
    Az[e] = acc;
  }
}

// ------------------------------------------------------------ 5. voxels
// One workgroup per image row (y, z).  B[dx][c2][c] = sum_dy Az[z][c][c2][dx][dy]
// exp(i theta_y dy) is built in LDS (contiguous in c); then a group of G2
// lanes (G2 >= NRx) per voxel, lane c holding row c of Cs(v) (FP32).  Power
// iteration in FP64, started from the previous voxel's eigenvector of the
// same lane group when that voxel was kept (else from the column of the
// largest diagonal entry), until the
// estimated distance to the limit, |dv| rho / (1 - rho) with rho the ratio
// of successive update norms, is below 1e-8, under the float32 rounding of
// the stored maps (or iters steps); a voxel
// whose trace or Frobenius norm already fails the mask (both bound
// lambda_max from above for a PSD matrix) skips the iteration.
constexpr int kVoxThreads = 256;
// B[dx][c2][c] of image row (y, z) into dst (LDS or global)
__device__ __forceinline__ void espirit_row_table(const float2* __restrict__ Az, int NRx, int kx, int ky, int Ny,
                                                  int y, int z, float2* dst) {
  const int Dx = 2 * kx - 1, Dy = 2 * ky - 1;
  const long long zoff = (long long)z * NRx * NRx * Dx * Dy;
  const int nb = NRx * NRx * Dx;
  for (int e = threadIdx.x; e < nb; e += blockDim.x) {
    const int dxi = e % Dx;
    const int cc = e / Dx;  // c * NRx + c2
    const int c = cc / NRx, c2 = cc % NRx;
    const float2* src = Az + zoff + (long long)e * Dy;
    float2 acc = make_float2(0.f, 0.f);
    for (int j = 0; j < Dy; ++j) acc = c_fma(src[j], esp_phase(y, Ny, j - (ky - 1)), acc);
    dst[(dxi * NRx + c2) * NRx + c] = acc;
  }
}

// Row tables of image rows row0 .. row0 + gridDim.x - 1 (row = y + Ny z) into
// global memory, for coil counts whose table does not fit LDS.
__global__ void __launch_bounds__(kVoxThreads) k_espirit_rows(const float2* __restrict__ Az, int NRx, int kx, int ky,
                                                              int Ny, int row0, float2* __restrict__ Bg) {
  const int row = row0 + blockIdx.x;
  espirit_row_table(Az, NRx, kx, ky, Ny, row % Ny, row / Ny,
                    Bg + (long long)blockIdx.x * (2 * kx - 1) * NRx * NRx);
}

// STAGED: the row table is built in LDS; else it is read from Bg (filled by
// k_espirit_rows for rows row0 + blockIdx.x).
template <int G2, bool STAGED>
__global__ void __launch_bounds__(kVoxThreads) k_espirit_voxels(const float2* __restrict__ Az, int NRx, int kx,
                                                                int ky, int Nx, int Ny, int Nz, float thr2k, int iters,
                                                                const float2* __restrict__ Bg, int row0,
                                                                float2* __restrict__ sens) {
  extern __shared__ double2 smem[];
  const int Dx = 2 * kx - 1;
  double2* vs = smem;  // kVoxThreads broadcast slots
  const int rowid = row0 + blockIdx.x;
  const int y = rowid % Ny, z = rowid / Ny;
  const float2* Bs;
  if (STAGED) {
    float2* Bl = (float2*)(smem + kVoxThreads);  // Dx * NRx * NRx
    espirit_row_table(Az, NRx, kx, ky, Ny, y, z, Bl);
    Bs = Bl;
  } else {
    Bs = Bg + (long long)blockIdx.x * Dx * NRx * NRx;
  }
  __syncthreads();
  const int lane = threadIdx.x % G2;
  const int grp = threadIdx.x / G2;
  constexpr int ngrp = kVoxThreads / G2;
  const bool act = lane < NRx;
  const int c = act ? lane : 0;
  double2* vg = vs + grp * G2;
  // lane groups interleave along the row (balanced across wavefronts); inside
  // the mask the maps are smooth, so a kept voxel's eigenvector starts the
  // group's next voxel (ngrp positions on) close to its limit, which does
  // not depend on the start; after a masked-out voxel the start is the
  // largest-diagonal column again
  double pvx = 0.0, pvy = 0.0;
  bool have_prev = false;  // group-uniform
  for (int x = grp; x < Nx; x += ngrp) {
    float2 row[G2];
#pragma unroll
    for (int c2 = 0; c2 < G2; ++c2) row[c2] = make_float2(0.f, 0.f);
    if (act) {
      // phases exp(i theta_x d), d = -(kx-1) .. kx-1: the first exactly, the
      // rest by the step exp(i theta_x) (a few float32 ulps over 2 kx - 2 steps)
      float2 ph = esp_phase(x, Nx, -(kx - 1));
      const float2 step = esp_phase(x, Nx, 1);
      for (int j = 0; j < Dx; ++j) {
        const float2* bj = Bs + (size_t)j * NRx * NRx + c;
#pragma unroll
        for (int c2 = 0; c2 < G2; ++c2)
          if (c2 < NRx) row[c2] = c_fma(bj[c2 * NRx], ph, row[c2]);
        ph = c_mul(ph, step);
      }
    }
    // diagonal (real) of this lane's row; trace and argmax over the group
    float dg = 0.f;
#pragma unroll
    for (int c2 = 0; c2 < G2; ++c2)
      if (c2 == c) dg = row[c2].x;
    float tr = dg;
#pragma unroll
    for (int o = G2 / 2; o > 0; o >>= 1) tr += __shfl_xor(tr, o, G2);
    float best = act ? dg : -1.f;
    int arg = lane;
#pragma unroll
    for (int o = G2 / 2; o > 0; o >>= 1) {
      const float ob = __shfl_xor(best, o, G2);
      const int oa = __shfl_xor(arg, o, G2);
      if (ob > best || (ob == best && oa < arg)) {
        best = ob;
        arg = oa;
      }
    }
    // ||Cs||_F >= lambda_max for the PSD Cs: below the threshold (less a
    // margin for FP32 rounding) the voxel cannot be kept either
    float fr = 0.f;
    if (act) {
#pragma unroll
      for (int c2 = 0; c2 < G2; ++c2) fr = fmaf(row[c2].x, row[c2].x, fmaf(row[c2].y, row[c2].y, fr));
    }
#pragma unroll
    for (int o = G2 / 2; o > 0; o >>= 1) fr += __shfl_xor(fr, o, G2);
    const bool frob_out = (double)fr <= (double)thr2k * (double)thr2k * (1.0 - 1e-4);
    float2 outv = make_float2(0.f, 0.f);
    if (tr > thr2k && !frob_out) {  // group-uniform
      // FP64 iterate (FP32 matrix): the update norm falls to ~1e-15, so the
      // error estimate dif * rho / (1 - rho) stays meaningful for slow rates
      double vx = 0.0, vy = 0.0;
#pragma unroll
      for (int c2 = 0; c2 < G2; ++c2)
        if (c2 == arg) {
          vx = row[c2].x;
          vy = row[c2].y;
        }
      if (have_prev) {
        vx = act ? pvx : 0.0;
        vy = act ? pvy : 0.0;
      }
      double nrm = vx * vx + vy * vy;
#pragma unroll
      for (int o = G2 / 2; o > 0; o >>= 1) nrm += __shfl_xor(nrm, o, G2);
      double inv = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
      vx *= inv;
      vy *= inv;
      // Each iterate is rescaled to unit norm with its coil-0 entry real and
      // non-negative: the FP32 coil matrix is Hermitian only to rounding, so
      // its dominant eigenvalue may carry a tiny phase that would otherwise
      // rotate the iterate every step and defeat the stopping test.
      double prev = 1.0, lam = 0.0;
      for (int it = 0; it < iters; ++it) {
        vg[lane] = make_double2(vx, vy);
        __builtin_amdgcn_wave_barrier();
        double wx = 0.0, wy = 0.0;
#pragma unroll
        for (int c2 = 0; c2 < G2; ++c2)
          if (c2 < NRx) {
            const double2 u = vg[c2];
            const double ax = row[c2].x, ay = row[c2].y;
            wx = fma(ax, u.x, fma(-ay, u.y, wx));
            wy = fma(ax, u.y, fma(ay, u.x, wy));
          }
        __builtin_amdgcn_wave_barrier();
        double wn = wx * wx + wy * wy;
        double rq = vx * wx + vy * wy;  // Re(v^H C v), v of unit norm
#pragma unroll
        for (int o = G2 / 2; o > 0; o >>= 1) {
          wn += __shfl_xor(wn, o, G2);
          rq += __shfl_xor(rq, o, G2);
        }
        lam = rq;
        const double w0x = __shfl(wx, 0, G2), w0y = __shfl(wy, 0, G2);
        const double a0 = sqrt(w0x * w0x + w0y * w0y);
        double px = 1.0, py = 0.0;
        if (a0 > 0.0) {
          px = w0x / a0;
          py = -w0y / a0;
        }
        inv = wn > 0.0 ? 1.0 / sqrt(wn) : 0.0;
        const double nx = (wx * px - wy * py) * inv, ny = (wx * py + wy * px) * inv;
        double dif = (nx - vx) * (nx - vx) + (ny - vy) * (ny - vy);
#pragma unroll
        for (int o = G2 / 2; o > 0; o >>= 1) dif += __shfl_xor(dif, o, G2);
        vx = nx;
        vy = ny;
        dif = sqrt(dif);
        const double rho = fmin(dif / fmax(prev, 1e-300), 0.999999);
        prev = dif;
        if (dif * rho <= 1e-8 * (1.0 - rho) || dif <= 1e-14) break;  // group-uniform
      }
      // lam: Rayleigh quotient of the last iterate, lambda_max(Cs)
      const bool valid = lam > (double)thr2k;  // group-uniform
      if (valid) outv = make_float2((float)vx, (float)vy);
      pvx = vx;
      pvy = vy;
      have_prev = valid;
    }
    if (act) sens[(((long long)x * Ny + y) * Nz + z) * NRx + lane] = outv;
  }
}

constexpr int kRowsPerChunk = 512;  // row tables held in global memory at once

template <int G2>
int launch_voxels(const float2* Az, int NRx, int kx, int ky, const int* imsize, float thr2k, int iters, float2* Bg,
                  float2* sens, hipStream_t stream) {
  const long long rows = (long long)imsize[1] * imsize[2];
  const size_t table = (size_t)NRx * NRx * (2 * kx - 1) * sizeof(float2);
  const size_t slots = kVoxThreads * sizeof(double2);
  if (table + slots <= 160 * 1024) {
    const size_t shmem = table + slots;
    ME_TRY_HIP(hipFuncSetAttribute((const void*)k_espirit_voxels<G2, true>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    hipLaunchKernelGGL((k_espirit_voxels<G2, true>), dim3((unsigned)rows), dim3(kVoxThreads), shmem, stream, Az, NRx,
                       kx, ky, imsize[0], imsize[1], imsize[2], thr2k, iters, (const float2*)nullptr, 0, sens);
  } else {
    ME_REQUIRE(Bg, "espirit_maps: workspace lacks the row tables for this coil count");
    for (long long r0 = 0; r0 < rows; r0 += kRowsPerChunk) {
      const int nr = (int)(rows - r0 < kRowsPerChunk ? rows - r0 : kRowsPerChunk);
      hipLaunchKernelGGL(k_espirit_rows, dim3((unsigned)nr), dim3(kVoxThreads), 0, stream, Az, NRx, kx, ky,
                         imsize[1], (int)r0, Bg);
      hipLaunchKernelGGL((k_espirit_voxels<G2, false>), dim3((unsigned)nr), dim3(kVoxThreads), slots, stream, Az,
                         NRx, kx, ky, imsize[0], imsize[1], imsize[2], thr2k, iters, (const float2*)Bg, (int)r0,
                         sens);
    }
  }
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

inline unsigned grid_for(long long n, int threads = 256, long long cap = 65536) {
  long long b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  return (unsigned)(b < cap ? b : cap);
}


// ------------------------------------------------------------ subspace tools
// C (m x n) = op(A) op(B), FP64 complex interleaved, row-major; op 0: as
// stored, op 1: conjugate transpose (A stored k x m, B stored n x k).
// 32 x 32 output tiles, 16-deep k tiles in LDS, 2 x 2 outputs per thread.
constexpr int kZT = 32, kZK = 16;
__global__ void __launch_bounds__(256) k_zgemm(int opa, int opb, int m, int n, int k, const double2* __restrict__ A,
                                               const double2* __restrict__ B, double2* __restrict__ Cm) {
  __shared__ double2 As[kZK][kZT + 1];
  __shared__ double2 Bs[kZK][kZT + 1];
  const int i0 = blockIdx.y * kZT, j0 = blockIdx.x * kZT;
  const int tid = threadIdx.x, ty = tid / 16, tx = tid % 16;
  double cr[2][2] = {{0, 0}, {0, 0}}, ci[2][2] = {{0, 0}, {0, 0}};
  for (int k0 = 0; k0 < k; k0 += kZK) {
    __syncthreads();
    for (int e = tid; e < kZT * kZK; e += 256) {
      int i, kk;
      double2 v = make_double2(0.0, 0.0);
      if (opa == 0) {  // A[i][kk], contiguous in kk
        i = e / kZK;
        kk = e % kZK;
        if (i0 + i < m && k0 + kk < k) v = A[(long long)(i0 + i) * k + k0 + kk];
      } else {  // conj(A[kk][i]), contiguous in i
        kk = e / kZT;
        i = e % kZT;
        if (i0 + i < m && k0 + kk < k) {
          v = A[(long long)(k0 + kk) * m + i0 + i];
          v.y = -v.y;
        }
      }
      As[kk][i] = v;
      int j;
      double2 w = make_double2(0.0, 0.0);
      if (opb == 0) {  // B[kk][j], contiguous in j
        kk = e / kZT;
        j = e % kZT;
        if (j0 + j < n && k0 + kk < k) w = B[(long long)(k0 + kk) * n + j0 + j];
      } else {  // conj(B[j][kk]), contiguous in kk
        j = e / kZK;
        kk = e % kZK;
        if (j0 + j < n && k0 + kk < k) {
          w = B[(long long)(j0 + j) * k + k0 + kk];
          w.y = -w.y;
        }
      }
      Bs[kk][j] = w;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < kZK; ++kk) {
      const double2 a0 = As[kk][2 * ty], a1 = As[kk][2 * ty + 1];
      const double2 b0 = Bs[kk][2 * tx], b1 = Bs[kk][2 * tx + 1];
      const double2 av[2] = {a0, a1}, bv[2] = {b0, b1};
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          cr[p][q] = fma(av[p].x, bv[q].x, fma(-av[p].y, bv[q].y, cr[p][q]));
          ci[p][q] = fma(av[p].x, bv[q].y, fma(av[p].y, bv[q].x, ci[p][q]));
        }
    }
  }
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int i = i0 + 2 * ty + p, j = j0 + 2 * tx + q;
      if (i < m && j < n) Cm[(long long)i * n + j] = make_double2(cr[p][q], ci[p][q]);
    }
}

// x[i] = complex standard normal from a counter-based hash of (seed, i):
// splitmix64 words, Box-Muller in FP64
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__global__ void k_random_normal(double2* __restrict__ x, long long n, unsigned long long seed) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const unsigned long long a = splitmix64(seed ^ (2ull * (unsigned long long)i));
    const unsigned long long b = splitmix64(seed ^ (2ull * (unsigned long long)i + 1ull));
    const double u1 = ((double)(a >> 11) + 0.5) * (1.0 / 9007199254740992.0);  // (0, 1)
    const double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
    const double r = sqrt(-2.0 * log(u1));
    double sn, cs;
    sincos(2.0 * kPi * u2, &sn, &cs);
    x[i] = make_double2(r * cs, r * sn);
  }
}

// X (rows x cols) *= diag(s) from the right: column j scaled by s[j]
__global__ void k_scale_columns(double2* __restrict__ X, long long rows, int cols, const double* __restrict__ s) {
  const long long n = rows * cols;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    const double f = s[e % cols];
    X[e].x *= f;
    X[e].y *= f;
  }
}

}  // namespace me

using namespace me;

extern "C" {

long long me_espirit_hankel_size(const int* calib_dims, const int* kernel) {
  if (!calib_dims || !kernel) return 0;
  long long nb = 1;
  for (int a = 0; a < 3; ++a) {
    if (kernel[a] < 1 || kernel[a] > calib_dims[a]) return 0;
    nb *= calib_dims[a] - kernel[a] + 1;
  }
  return (long long)kernel[0] * kernel[1] * kernel[2] * calib_dims[3] * nb;
}

int me_espirit_calibration_matrix(const void* calib, const int* calib_dims, const int* kernel, void* H,
                                  hipStream_t stream) {
  ME_REQUIRE(calib && calib_dims && kernel && H, "espirit_calibration_matrix: null argument");
  const long long total = me_espirit_hankel_size(calib_dims, kernel);
  ME_REQUIRE(total > 0, "espirit_calibration_matrix: kernel must fit inside the calibration region");
  hipLaunchKernelGGL(k_hankel, dim3(grid_for(total)), dim3(256), 0, stream, (const float2*)calib, calib_dims[0],
                     calib_dims[1], calib_dims[2], calib_dims[3], kernel[0], kernel[1], kernel[2], (float2*)H);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_gram_rows(const void* X, int m, long long n, double* G, hipStream_t stream) {
  ME_REQUIRE(X && G && m >= 1 && n >= 1, "gram_rows: bad argument");
  const long long nt = (m + kGT - 1) / kGT;
  const long long pairs = nt * (nt + 1) / 2;
  ME_REQUIRE(pairs < (1LL << 31), "gram_rows: too many tiles");
  hipLaunchKernelGGL(k_gram_tiled, dim3((unsigned)pairs), dim3(256), 0, stream, (const float2*)X, m, n, G);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

long long me_jacobi_workspace_size(int m) {
  return (long long)kRoundsPerVPass * ((m + 2) / 2) * 4 + 2 * kNormBlocks;
}

int me_jacobi_set_lds_rows(int rows) {
  return g_jac_lds_rows.exchange(rows < 0 ? 0 : (rows > kMaxLdsRow ? kMaxLdsRow : rows));
}

int me_jacobi_eig(double* G, double* V, int m, int max_sweeps, double tol, double* eigvals, double* work,
                  int* sweeps_done, hipStream_t stream) {
  ME_REQUIRE(G && V && eigvals && work && m >= 1, "jacobi_eig: bad argument");
  ME_REQUIRE(m <= 32768, "jacobi_eig: at most 32768 rows");
  const int mm = m + (m & 1);
  const int npairs = mm / 2;
  const bool lds_v = m <= g_jac_lds_rows.load();
  double* rots = work;  // kRoundsPerVPass rounds of npairs x 4
  double* part = work + (long long)kRoundsPerVPass * npairs * 4;
  const size_t vshm = (size_t)m * sizeof(double2);
  if (lds_v)
    ME_TRY_HIP(hipFuncSetAttribute((const void*)k_jac_apply_v, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)vshm));
  hipLaunchKernelGGL(k_identity, dim3(grid_for((long long)m * m)), dim3(256), 0, stream, V, m);
  ME_TRY_HIP(hipGetLastError());
  int sw = 0;
  std::vector<double> hp(2 * kNormBlocks);
  for (; m >= 2 && sw < max_sweeps; ++sw) {
    hipLaunchKernelGGL(k_jac_norms, dim3(kNormBlocks), dim3(256), 0, stream, (const double*)G, m, part);
    ME_TRY_HIP(hipGetLastError());
    ME_TRY_HIP(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    ME_TRY_HIP(hipStreamSynchronize(stream));
    double off = 0.0, dia = 0.0;
    for (int b = 0; b < kNormBlocks; ++b) {
      off += hp[2 * b];
      dia += hp[2 * b + 1];
    }
    if (off <= tol * tol * dia) break;
    const dim3 grid((unsigned)((npairs + 255) / 256), (unsigned)(npairs + (lds_v ? 0 : m)));
    for (int r0 = 0; r0 < mm - 1; r0 += kRoundsPerVPass) {
      const int nr = (mm - 1 - r0) < kRoundsPerVPass ? (mm - 1 - r0) : kRoundsPerVPass;
      for (int k = 0; k < nr; ++k) {
        double* rot = rots + (long long)k * npairs * 4;
        hipLaunchKernelGGL(k_jac_params, dim3((npairs + 255) / 256), dim3(256), 0, stream, (const double*)G, m, mm,
                           r0 + k, rot);
        hipLaunchKernelGGL(k_jac_round, grid, dim3(256), 0, stream, G, V, m, mm, r0 + k, (const double*)rot);
      }
      if (lds_v)
        hipLaunchKernelGGL(k_jac_apply_v, dim3((unsigned)m), dim3(256), vshm, stream, V, m, mm, r0, nr,
                           (const double*)rots);
    }
    ME_TRY_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_diag, dim3((m + 255) / 256), dim3(256), 0, stream, (const double*)G, m, eigvals);
  ME_TRY_HIP(hipGetLastError());
  if (sweeps_done) *sweeps_done = sw;
  return 0;
}

int me_zgemm(int opa, int opb, int m, int n, int k, const double* A, const double* B, double* C, hipStream_t stream) {
  ME_REQUIRE(A && B && C && m >= 1 && n >= 1 && k >= 1, "zgemm: bad argument");
  ME_REQUIRE((opa == 0 || opa == 1) && (opb == 0 || opb == 1), "zgemm: op must be 0 (as stored) or 1 (conj transpose)");
  ME_REQUIRE((m + kZT - 1) / kZT < 65536, "zgemm: too many rows");
  dim3 grid((unsigned)((n + kZT - 1) / kZT), (unsigned)((m + kZT - 1) / kZT));
  hipLaunchKernelGGL(k_zgemm, grid, dim3(256), 0, stream, opa, opb, m, n, k, (const double2*)A, (const double2*)B,
                     (double2*)C);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_random_normal(double* x, long long n, unsigned long long seed, hipStream_t stream) {
  ME_REQUIRE(x && n >= 1, "random_normal: bad argument");
  hipLaunchKernelGGL(k_random_normal, dim3(grid_for(n)), dim3(256), 0, stream, (double2*)x, n, seed);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_scale_columns(double* X, long long rows, int cols, const double* s, hipStream_t stream) {
  ME_REQUIRE(X && s && rows >= 1 && cols >= 1, "scale_columns: bad argument");
  hipLaunchKernelGGL(k_scale_columns, dim3(grid_for(rows * cols)), dim3(256), 0, stream, (double2*)X, rows, cols, s);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_espirit_projector(const double* V, int m, int ldv, const int* idx, int r, void* Xwork, double* P,
                         hipStream_t stream) {
  ME_REQUIRE(V && idx && Xwork && P && m >= 1 && r >= 1 && r <= ldv, "espirit_projector: bad argument");
  hipLaunchKernelGGL(k_gather_columns, dim3(grid_for((long long)m * r)), dim3(256), 0, stream, V, m, ldv, idx, r,
                     (float2*)Xwork);
  ME_TRY_HIP(hipGetLastError());
  return me_gram_rows(Xwork, m, r, P, stream);
}

long long me_espirit_maps_workspace_size(int NRx, const int* kernel, const int* imsize) {
  if (!kernel || !imsize || NRx < 1) return 0;
  const long long dxy = (long long)(2 * kernel[0] - 1) * (2 * kernel[1] - 1);
  long long n = (long long)NRx * NRx * dxy * ((2 * kernel[2] - 1) + imsize[2]);  // complex64 elements: Q, Az
  const long long table = (long long)NRx * NRx * (2 * kernel[0] - 1);
  if (table * 8 + kVoxThreads * 16 > 160 * 1024) n += table * kRowsPerChunk;  // row tables in global memory
  return n;
}

int me_espirit_maps(const double* P, int NRx, const int* kernel, const int* imsize, double mask_thresh, int iters,
                    void* work, void* sens, hipStream_t stream) {
  ME_REQUIRE(P && kernel && imsize && work && sens, "espirit_maps: null argument");
  ME_REQUIRE(NRx >= 1 && NRx <= 64, "espirit_maps: 1 to 64 coils");
  ME_REQUIRE(iters >= 1, "espirit_maps: iters must be positive");
  for (int a = 0; a < 3; ++a) ME_REQUIRE(kernel[a] >= 1 && imsize[a] >= 1, "espirit_maps: bad sizes");
  const int kx = kernel[0], ky = kernel[1], kz = kernel[2];
  float2* Q = (float2*)work;
  const long long nq = (long long)NRx * NRx * (2 * kx - 1) * (2 * ky - 1) * (2 * kz - 1);
  float2* Az = Q + nq;
  hipLaunchKernelGGL(k_espirit_q, dim3(grid_for(nq)), dim3(256), 0, stream, P, NRx, kx, ky, kz, Q);
  const long long naz = (long long)NRx * NRx * (2 * kx - 1) * (2 * ky - 1) * imsize[2];
  hipLaunchKernelGGL(k_espirit_az, dim3(grid_for(naz)), dim3(256), 0, stream, (const float2*)Q, NRx, kx, ky, kz,
                     imsize[2], Az);
  ME_TRY_HIP(hipGetLastError());
  float2* Bg = Az + naz;  // row tables, used when they exceed LDS
  ME_REQUIRE((long long)imsize[1] * imsize[2] < (1LL << 31), "espirit_maps: image too large");
  // mask: sqrt(lambda(Cs)) / sqrt(prod kernel) > mask_thresh
  // (a negative threshold passes every voxel with a nonzero principal vector)
  const float thr2k = mask_thresh < 0.0 ? -1.f : (float)(mask_thresh * mask_thresh * (double)kx * ky * kz);
  if (NRx <= 8) return launch_voxels<8>(Az, NRx, kx, ky, imsize, thr2k, iters, Bg, (float2*)sens, stream);
  if (NRx <= 16) return launch_voxels<16>(Az, NRx, kx, ky, imsize, thr2k, iters, Bg, (float2*)sens, stream);
  if (NRx <= 32) return launch_voxels<32>(Az, NRx, kx, ky, imsize, thr2k, iters, Bg, (float2*)sens, stream);
  return launch_voxels<64>(Az, NRx, kx, ky, imsize, thr2k, iters, Bg, (float2*)sens, stream);
}

}  // extern "C"
