// FFT-shear rigid transform T(theta), its adjoint and the per-parameter
// derivative operator, composed from line passes (linepass.hip).
//
//  apply   (transform.py:263-318):  T = Tr(q) Rz(t3) Ry(t2) Rx(t1), each R three
//          1-D Fourier shears tan/sin/tan.  Translation is separable into one
//          phase per axis: Tr_x folds into Rz's last axis-0 pass, Tr_z (which
//          commutes with Rz) into Ry's last axis-2 pass, Tr_y is a 10th pass.
//  adjoint (transform.py:320-377):  conjugate translation first, then Rz^H,
//          Ry^H, Rx^H with conjugated shears; Tr_x^H and Tr_z^H fold the same way.
//  deriv   (transform.py:379-504):  d/dp of T_d = Tr Rx Ry Rz (rotations z,y,x).
//          Rotation derivatives use the three-term product rule as three passes
//          with 1->2, 2->2, 2->1 buffers; translation derivatives multiply the
//          axis-p translation phase by -i k_p.
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "me_device.h"
#include "me_internal.h"
#include "../../include/motion_estimation_amd.h"

namespace me {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* get_error() { return g_err.c_str(); }

// Parameter row [q1,q2,q3,t1,t2,t3] -> coefficient row (FP64):
// [q1,q2,q3, tan(t1/2),sin t1,cos t1,(1+tan^2(t1/2))/2, ... for t2, t3, 0]
__global__ void k_rigid_coef(const double* __restrict__ theta, long long th_stride, double* __restrict__ coef,
                             int nbatch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbatch) return;
  const double* p = theta + (long long)b * th_stride;
  double* c = coef + (long long)b * kNumCoef;
  c[0] = p[0];
  c[1] = p[1];
  c[2] = p[2];
  for (int rho = 0; rho < 3; ++rho) {
    const double th = p[3 + rho];
    const double t = tan(0.5 * th);
    c[3 + 4 * rho] = t;
    c[4 + 4 * rho] = sin(th);
    c[5 + 4 * rho] = cos(th);
    c[6 + 4 * rho] = 0.5 * (1.0 + t * t);
  }
  c[15] = 0.0;
}

namespace {

constexpr int kTanAxis[3] = {1, 2, 0};  // rotation about x, y, z: tan-shear axis
constexpr int kSinAxis[3] = {2, 0, 1};  //                         sin-shear axis
constexpr int kTanB[3] = {2, 0, 1};     // rk_tan = r_b * k_a (utils.py:124-129)
constexpr int kSinB[3] = {1, 2, 0};     // rk_sin = r_b * k_a (utils.py:130-134)

HostPhase shear(int rho, bool is_tan, bool conj) {
  HostPhase h;
  h.active = true;
  if (is_tan) {  // exp(i tan(t/2) rk)   (transform.py:275)
    h.ic = ci_tan(rho);
    h.sc = 1.f;
    h.b_axis = kTanB[rho];
  } else {  // exp(-i sin(t) rk)        (transform.py:276)
    h.ic = ci_sin(rho);
    h.sc = -1.f;
    h.b_axis = kSinB[rho];
  }
  if (conj) h.sc = -h.sc;
  return h;
}
// times exp(-i q_j k_a) (transform.py:257-261), conjugated if conj
HostPhase with_trans(HostPhase h, int j, bool conj) {
  h.active = true;
  h.id = j;
  h.sd = conj ? 1.f : -1.f;
  return h;
}
HostPhase trans(int j, bool conj) { return with_trans(HostPhase{}, j, conj); }
// vtan' = i (1+tan^2(t/2))/2 rk vtan  (transform.py:437)
HostPhase with_dtan(HostPhase h, int rho) {
  h.deriv = 1;
  h.ig = ci_g(rho);
  h.sg = 1.f;
  h.h = 0.f;
  return h;
}
// vsin' = -i cos(t) rk vsin  (transform.py:438)
HostPhase with_dsin(HostPhase h, int rho) {
  h.deriv = 1;
  h.ig = ci_cos(rho);
  h.sg = -1.f;
  h.h = 0.f;
  return h;
}
// -i k_p U  (transform.py:428)
HostPhase with_dtrans(HostPhase h) {
  h.deriv = 1;
  h.ig = -1;
  h.sg = 0.f;
  h.h = -1.f;
  return h;
}

enum Buf { BX = 0, BY = 1, BW = 2 };
struct Pass {
  int axis = 0;
  int nin = 1, nout = 1;
  int in[2] = {BY, BY};
  int out[2] = {BY, BY};
  HostPhase ph[2][2];
};

Pass single(int axis, HostPhase p) {
  Pass q;
  q.axis = axis;
  q.ph[0][0] = p;
  return q;
}

void add_rotation(std::vector<Pass>& v, int rho, bool conj, int fold_j, bool fold_conj, bool fold_first,
                  bool dtrans_last) {
  const int at = kTanAxis[rho], as = kSinAxis[rho];
  HostPhase t1 = shear(rho, true, conj), s = shear(rho, false, conj), t2 = shear(rho, true, conj);
  if (fold_j >= 0) {
    if (fold_first)
      t1 = with_trans(t1, fold_j, fold_conj);
    else
      t2 = with_trans(t2, fold_j, fold_conj);
  }
  if (dtrans_last) t2 = with_dtrans(t2);
  v.push_back(single(at, t1));
  v.push_back(single(as, s));
  v.push_back(single(at, t2));
}

// Product rule for d/dtheta of R = D_t(vtan) D_s(vsin) D_t(vtan)
// (transform.py:433-474):
//   P1 (axis t): A = D_t(vtan') X,  B = D_t(vtan) X
//   P2 (axis s): A = D_s(vsin) A + D_s(vsin') B,  B = D_s(vsin) B
//   P3 (axis t): Y = D_t(vtan) A + D_t(vtan') B    (optionally times a translation)
void add_rotation_deriv(std::vector<Pass>& v, int rho, int fold_j) {
  const int at = kTanAxis[rho], as = kSinAxis[rho];
  Pass p1;
  p1.axis = at;
  p1.nin = 1;
  p1.nout = 2;
  p1.in[0] = BY;
  p1.out[0] = BY;
  p1.out[1] = BW;
  p1.ph[0][0] = with_dtan(shear(rho, true, false), rho);
  p1.ph[1][0] = shear(rho, true, false);
  v.push_back(p1);
  Pass p2;
  p2.axis = as;
  p2.nin = 2;
  p2.nout = 2;
  p2.in[0] = BY;
  p2.in[1] = BW;
  p2.out[0] = BY;
  p2.out[1] = BW;
  p2.ph[0][0] = shear(rho, false, false);
  p2.ph[0][1] = with_dsin(shear(rho, false, false), rho);
  p2.ph[1][1] = shear(rho, false, false);
  v.push_back(p2);
  Pass p3;
  p3.axis = at;
  p3.nin = 2;
  p3.nout = 1;
  p3.in[0] = BY;
  p3.in[1] = BW;
  p3.out[0] = BY;
  HostPhase a = shear(rho, true, false), b = with_dtan(shear(rho, true, false), rho);
  if (fold_j >= 0) {
    a = with_trans(a, fold_j, false);
    b = with_trans(b, fold_j, false);
  }
  p3.ph[0][0] = a;
  p3.ph[0][1] = b;
  v.push_back(p3);
}

std::vector<Pass> build_apply(bool inverse_translation) {
  // transform.py:267-318: x, y, z rotations then translation
  std::vector<Pass> v;
  const bool ct = inverse_translation;
  add_rotation(v, 0, false, -1, false, false, false);
  add_rotation(v, 1, false, 2, ct, false, false);  // Tr_z folded (commutes with Rz)
  add_rotation(v, 2, false, 0, ct, false, false);  // Tr_x folded
  v.push_back(single(1, trans(1, ct)));            // Tr_y
  return v;
}

std::vector<Pass> build_adjoint() {
  // transform.py:331-377: conjugate translation, then z, y, x conjugated
  std::vector<Pass> v;
  v.push_back(single(1, trans(1, true)));  // Tr_y^H
  add_rotation(v, 2, true, 0, true, true, false);  // Tr_x^H folded into Rz^H's first pass
  add_rotation(v, 1, true, 2, true, true, false);  // Tr_z^H (commutes with Rz^H) into Ry^H's first pass
  add_rotation(v, 0, true, -1, false, false, false);
  return v;
}

std::vector<Pass> build_deriv(int p) {
  // transform.py:476-504: rotations z, y, x (one of them differentiated for
  // p >= 3), then translation (differentiated along axis p for p < 3).
  std::vector<Pass> v;
  if (p == 5)
    add_rotation_deriv(v, 2, -1);
  else
    add_rotation(v, 2, false, -1, false, false, false);
  if (p == 4)
    add_rotation_deriv(v, 1, -1);
  else
    add_rotation(v, 1, false, -1, false, false, false);
  if (p == 3)
    add_rotation_deriv(v, 0, 1);  // Tr_y folded into the last axis-1 pass
  else
    add_rotation(v, 0, false, 1, false, false, p == 1);
  HostPhase tx = trans(0, false), tz = trans(2, false);
  if (p == 0) tx = with_dtrans(tx);
  if (p == 2) tz = with_dtrans(tz);
  v.push_back(single(0, tx));
  v.push_back(single(2, tz));
  return v;
}

// Shared-chain derivative columns.  With T_d = Tr Rx Ry Rz, the six columns
// need the rotation chain up to different levels:
//   c1 = Rz x,  c2 = Ry c1,  c3 = Tr Rx c2 = T_d x (Tr_y folded into Rx's
//   last pass, then Tr_x and Tr_z passes)
//   p = 0,1,2: D_p c3              (one pass along axis p with the factor -i k_p;
//                                   D_p commutes with the translations)
//   p = 3:     Tr dRx c2           (Tr_y folded into dRx's last pass, then Tr_x, Tr_z)
//   p = 4:     Tr Rx dRy c1
//   p = 5:     Tr Rx Ry dRz x
// 11 + 3*1 + 5 + 8 + 11 = 38 passes for all six columns instead of 6*11.
std::vector<Pass> build_rotation_only(int rho) {
  std::vector<Pass> v;
  add_rotation(v, rho, false, -1, false, false, false);
  return v;
}

// Last level of the shared chain: c3 = Tr Rx c2 (the whole T_d x).
std::vector<Pass> build_chain_last() {
  std::vector<Pass> v;
  add_rotation(v, 0, false, 1, false, false, false);  // Tr_y folded into the last axis-1 pass
  v.push_back(single(0, trans(0, false)));
  v.push_back(single(2, trans(2, false)));
  return v;
}

std::vector<Pass> build_deriv_from_chain(int p) {
  std::vector<Pass> v;
  if (p <= 2) {  // -i k_p along axis p of c3 = T_d x
    HostPhase d;
    d.active = true;
    v.push_back(single(p, with_dtrans(d)));
    return v;
  }
  if (p == 5) {
    add_rotation_deriv(v, 2, -1);
  }
  if (p >= 4) {
    if (p == 4)
      add_rotation_deriv(v, 1, -1);
    else
      add_rotation(v, 1, false, -1, false, false, false);
  }
  if (p == 3)
    add_rotation_deriv(v, 0, 1);
  else
    add_rotation(v, 0, false, 1, false, false, false);
  v.push_back(single(0, trans(0, false)));
  v.push_back(single(2, trans(2, false)));
  return v;
}

// Jacobian columns of the apply order T = Tr Rz Ry Rx (transform.py:267-318),
// for gradients through ``apply``.  With c1 = Rx x and c2 = Tr_z Ry c1 (Tr_z
// folded into Ry's last pass, as in apply) and u = T x:
//   p = 0,1,2: D_p u, one pass along axis p with the factor -i k_p
//   p = 3:     Tr_y (Rz Tr_x) (Ry Tr_z) dRx x      (10 passes)
//   p = 4:     Tr_y (Rz Tr_x) d(Tr_z Ry) c1        (7 passes)
//   p = 5:     Tr_y d(Tr_x Rz) c2                  (4 passes)
// so the six columns take 3 + 3 + 3 + 10 + 7 + 4 = 30 passes given u.
// level 0: c1 = Rx x; 1: c2 = Tr_z Ry c1; 2: u = Tr_y (Rz Tr_x) c2 = T x.
// The three levels are build_apply(false)'s passes in its order, so u equals
// me_rigid_transform's apply bit for bit.
std::vector<Pass> build_apply_chain(int level) {
  std::vector<Pass> v;
  if (level == 0) {
    add_rotation(v, 0, false, -1, false, false, false);
  } else if (level == 1) {
    add_rotation(v, 1, false, 2, false, false, false);
  } else {
    add_rotation(v, 2, false, 0, false, false, false);
    v.push_back(single(1, trans(1, false)));
  }
  return v;
}

std::vector<Pass> build_deriv_apply(int p) {
  std::vector<Pass> v;
  if (p <= 2) {
    HostPhase d;
    d.active = true;
    v.push_back(single(p, with_dtrans(d)));
    return v;
  }
  if (p == 3) {
    add_rotation_deriv(v, 0, -1);
    add_rotation(v, 1, false, 2, false, false, false);
  } else if (p == 4) {
    add_rotation_deriv(v, 1, 2);
  }
  if (p == 5)
    add_rotation_deriv(v, 2, 0);
  else
    add_rotation(v, 2, false, 0, false, false, false);
  v.push_back(single(1, trans(1, false)));
  return v;
}

struct Bufs {
  const float2* x;
  long long x_bs;
  int x_parts;  // partial volumes per batch entry summed by the first pass (1 or 2)
  float2* y;
  long long y_bs;
  float2* w;
  long long w_bs;
};

// Dot epilogue of the last pass (see LineArgs::dot_w): its output is not
// stored; the per-workgroup partials go to `partial` and their count per
// batch entry to *grid_x.
struct DotOut {
  const float2* w = nullptr;
  long long w_bs = 0;
  double* partial = nullptr;
  long long* grid_x = nullptr;
};

int run_passes(const std::vector<Pass>& passes, const VolGeom& g, Bufs b, int nbatch, const double* coef,
               float2* sum_out, const unsigned char* mask, bool accumulate, hipStream_t stream,
               const DotOut& dot = DotOut{}) {
  for (size_t i = 0; i < passes.size(); ++i) {
    const Pass& p = passes[i];
    LineArgs A{};
    A.nbatch = nbatch;
    A.coef = coef;
    A.ncoef = kNumCoef;
    A.mode = 0;
    int nin = p.nin;
    HostPhase ph[2][2];
    for (int o = 0; o < 2; ++o)
      for (int k = 0; k < 2; ++k) ph[o][k] = p.ph[o][k];
    for (int k = 0; k < p.nin; ++k) {
      int bi = (i == 0 && p.in[k] == BY) ? BX : p.in[k];
      A.in[k] = bi == BX ? b.x : (bi == BY ? b.y : b.w);
      A.in_bs[k] = bi == BX ? b.x_bs : (bi == BY ? b.y_bs : b.w_bs);
      ME_REQUIRE(A.in[k] != nullptr, "rigid: missing buffer");
    }
    if (i == 0 && b.x_parts == 2) {  // sum two partial inputs inside the first pass
      ME_REQUIRE(p.nin == 1 && p.in[0] == BY, "rigid: partial inputs need a single-input first pass");
      nin = 2;
      A.in[1] = b.x + (long long)g.n[0] * g.n[1] * g.n[2];  // next part volume
      A.in_bs[1] = b.x_bs;
      for (int o = 0; o < p.nout; ++o) ph[o][1] = ph[o][0];
    }
    for (int k = 0; k < p.nout; ++k) {
      A.out[k] = p.out[k] == BY ? b.y : b.w;
      A.out_bs[k] = p.out[k] == BY ? b.y_bs : b.w_bs;
      ME_REQUIRE(A.out[k] != nullptr, "rigid: missing work buffer");
    }
    const bool last = i + 1 == passes.size();
    if (last && sum_out) {
      A.reduce = 1;
      A.out[0] = sum_out;
      A.out_bs[0] = 0;
      A.mask = mask;
      A.mask_bs = 0;
      A.accumulate = accumulate ? 1 : 0;
    }
    long long* gx = nullptr;
    if (last && dot.w) {
      A.dot_w = dot.w;
      A.dot_w_bs = dot.w_bs;
      A.dot_partial = dot.partial;
      gx = dot.grid_x;
    }
    int rc = launch_linepass(g, p.axis, nin, p.nout, A, ph, stream, gx);
    if (rc) return rc;
  }
  return 0;
}

// out[b * out_stride] = sum of the nblk partials of batch entry b, in order.
__global__ void k_dot_finalize(const double* __restrict__ partial, long long nblk, double* __restrict__ out,
                               int out_stride) {
  __shared__ double sh[4];
  const int b = blockIdx.x;
  double v = 0.0;
  for (long long i = threadIdx.x; i < nblk; i += blockDim.x) v += partial[(long long)b * nblk + i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) out[(long long)b * out_stride] = sh[0] + sh[1] + sh[2] + sh[3];
}

int make_geom(VolGeom& g, const int* n, const long long* st, const double* r0, const double* r1,
              const double* r2) {
  ME_REQUIRE(n && st, "null shape/stride");
  for (int a = 0; a < 3; ++a) {
    ME_REQUIRE(n[a] > 0, "non-positive extent");
    g.n[a] = n[a];
    g.st[a] = st[a];
  }
  g.r[0] = r0;
  g.r[1] = r1;
  g.r[2] = r2;
  return 0;
}

int launch_coef(const double* theta, long long th_stride, double* coef, int nbatch, hipStream_t s) {
  ME_REQUIRE(theta && coef, "rigid: null parameter or coefficient buffer");
  hipLaunchKernelGGL(k_rigid_coef, dim3((nbatch + 63) / 64), dim3(64), 0, s, theta, th_stride, coef, nbatch);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

}  // namespace
}  // namespace me

using namespace me;

extern "C" {

const char* me_last_error(void) { return me::get_error(); }

int me_version(void) { return ME_ABI_VERSION; }

int me_rigid_transform(const void* x, long long x_bstride, int x_parts, void* y, long long y_bstride, void* work,
                       long long work_bstride, int nbatch, const int* n, const long long* st, const double* r0,
                       const double* r1, const double* r2, const double* theta, long long theta_stride,
                       double* coef_ws, int mode, int flags, const unsigned char* mask, hipStream_t stream) {
  VolGeom g;
  if (int rc = make_geom(g, n, st, r0, r1, r2)) return rc;
  ME_REQUIRE(x && y, "rigid_transform: null input/output");
  ME_REQUIRE(x_parts == 1 || x_parts == 2, "rigid_transform: x_parts must be 1 or 2");
  ME_REQUIRE(nbatch >= 1, "rigid_transform: nbatch < 1");
  ME_REQUIRE(mode >= 0 && mode <= 2, "rigid_transform: mode must be 0 (apply), 1 (adjoint), 2 (apply, inverse translation)");
  const bool sum = flags & ME_FLAG_SUM_BATCH;
  ME_REQUIRE(!sum || work, "rigid_transform: batch sum needs a work buffer");
  ME_REQUIRE(!(flags & ME_FLAG_MASK) || mask, "rigid_transform: mask flag without mask");
  ME_REQUIRE(!(flags & ME_FLAG_ACCUMULATE) || sum, "rigid_transform: accumulate needs the batch sum");
  if (int rc = launch_coef(theta, theta_stride, coef_ws, nbatch, stream)) return rc;
  std::vector<Pass> passes = mode == 1 ? build_adjoint() : build_apply(mode == 2);
  Bufs b{(const float2*)x, x_bstride, x_parts, (float2*)(sum ? work : y), sum ? work_bstride : y_bstride, nullptr, 0};
  return run_passes(passes, g, b, nbatch, coef_ws, sum ? (float2*)y : nullptr,
                    (flags & ME_FLAG_MASK) ? mask : nullptr, (flags & ME_FLAG_ACCUMULATE) != 0, stream);
}

int me_rigid_deriv(const void* x, long long x_bstride, void* y, long long y_bstride, void* work,
                   long long work_bstride, int nbatch, const int* n, const long long* st, const double* r0,
                   const double* r1, const double* r2, const double* theta, long long theta_stride,
                   double* coef_ws, int p_idx, hipStream_t stream) {
  VolGeom g;
  if (int rc = make_geom(g, n, st, r0, r1, r2)) return rc;
  ME_REQUIRE(x && y, "rigid_deriv: null input/output");
  ME_REQUIRE(p_idx >= 0 && p_idx <= 5, "rigid_deriv: p_idx must be in [0, 5]");
  ME_REQUIRE(p_idx < 3 || work, "rigid_deriv: rotation derivatives need a work buffer");
  if (int rc = launch_coef(theta, theta_stride, coef_ws, nbatch, stream)) return rc;
  Bufs b{(const float2*)x, x_bstride, 1, (float2*)y, y_bstride, (float2*)work, work_bstride};
  return run_passes(build_deriv(p_idx), g, b, nbatch, coef_ws, nullptr, nullptr, false, stream);
}

int me_rigid_deriv_chain(const void* x, long long x_bstride, void* c1, void* c2, void* c3, long long c_bstride,
                         void* y, long long y_bstride, void* work, long long work_bstride, int nbatch, const int* n,
                         const long long* st, const double* r0, const double* r1, const double* r2,
                         const double* theta, long long theta_stride, double* coef_ws, int p_idx, int build,
                         hipStream_t stream) {
  VolGeom g;
  if (int rc = make_geom(g, n, st, r0, r1, r2)) return rc;
  ME_REQUIRE(x && c1 && c2 && c3 && y, "rigid_deriv_chain: null buffer");
  ME_REQUIRE(p_idx >= 0 && p_idx <= 5, "rigid_deriv_chain: p_idx must be in [0, 5]");
  ME_REQUIRE(p_idx < 3 || work, "rigid_deriv_chain: rotation derivatives need a work buffer");
  if (int rc = launch_coef(theta, theta_stride, coef_ws, nbatch, stream)) return rc;
  if (build) {
    const void* src[3] = {x, c1, c2};
    const long long sbs[3] = {x_bstride, c_bstride, c_bstride};
    void* dst[3] = {c1, c2, c3};
    const int rho[2] = {2, 1};
    for (int l = 0; l < 3; ++l) {
      Bufs b{(const float2*)src[l], sbs[l], 1, (float2*)dst[l], c_bstride, nullptr, 0};
      const std::vector<Pass> passes = l < 2 ? build_rotation_only(rho[l]) : build_chain_last();
      if (int rc = run_passes(passes, g, b, nbatch, coef_ws, nullptr, nullptr, false, stream)) return rc;
    }
  }
  const void* from = p_idx <= 2 ? c3 : (p_idx == 3 ? c2 : (p_idx == 4 ? c1 : x));
  const long long fbs = p_idx == 5 ? x_bstride : c_bstride;
  Bufs b{(const float2*)from, fbs, 1, (float2*)y, y_bstride, (float2*)work, work_bstride};
  return run_passes(build_deriv_from_chain(p_idx), g, b, nbatch, coef_ws, nullptr, nullptr, false, stream);
}

int me_rigid_deriv_apply(const void* x, long long x_bstride, const void* u, long long u_bstride, void* c1, void* c2,
                         long long c_bstride, void* y, long long y_bstride, void* work, long long work_bstride,
                         int nbatch, const int* n, const long long* st, const double* r0, const double* r1,
                         const double* r2, const double* theta, long long theta_stride, double* coef_ws, int p_idx,
                         int build, hipStream_t stream) {
  VolGeom g;
  if (int rc = make_geom(g, n, st, r0, r1, r2)) return rc;
  ME_REQUIRE(y, "rigid_deriv_apply: null output");
  ME_REQUIRE(p_idx >= 0 && p_idx <= 5, "rigid_deriv_apply: p_idx must be in [0, 5]");
  ME_REQUIRE(p_idx >= 3 || u, "rigid_deriv_apply: translation columns need u = T x");
  ME_REQUIRE(p_idx < 3 || work, "rigid_deriv_apply: rotation columns need a work buffer");
  ME_REQUIRE(!build || (x && c1 && c2), "rigid_deriv_apply: building the chain needs x, c1 and c2");
  ME_REQUIRE(p_idx != 3 || x, "rigid_deriv_apply: column 3 needs x");
  ME_REQUIRE(p_idx != 4 || c1, "rigid_deriv_apply: column 4 needs c1");
  ME_REQUIRE(p_idx != 5 || c2, "rigid_deriv_apply: column 5 needs c2");
  if (int rc = launch_coef(theta, theta_stride, coef_ws, nbatch, stream)) return rc;
  if (build) {
    Bufs b0{(const float2*)x, x_bstride, 1, (float2*)c1, c_bstride, nullptr, 0};
    if (int rc = run_passes(build_apply_chain(0), g, b0, nbatch, coef_ws, nullptr, nullptr, false, stream)) return rc;
    Bufs b1{(const float2*)c1, c_bstride, 1, (float2*)c2, c_bstride, nullptr, 0};
    if (int rc = run_passes(build_apply_chain(1), g, b1, nbatch, coef_ws, nullptr, nullptr, false, stream)) return rc;
  }
  const void* from = p_idx <= 2 ? u : (p_idx == 3 ? x : (p_idx == 4 ? (const void*)c1 : (const void*)c2));
  const long long fbs = p_idx <= 2 ? u_bstride : (p_idx == 3 ? x_bstride : c_bstride);
  Bufs b{(const float2*)from, fbs, 1, (float2*)y, y_bstride, (float2*)work, work_bstride};
  return run_passes(build_deriv_apply(p_idx), g, b, nbatch, coef_ws, nullptr, nullptr, false, stream);
}

int me_rigid_apply_chain(const void* x, long long x_bstride, void* c1, void* c2, long long c_bstride, void* u,
                         long long u_bstride, int nbatch, const int* n, const long long* st, const double* r0,
                         const double* r1, const double* r2, const double* theta, long long theta_stride,
                         double* coef_ws, hipStream_t stream) {
  VolGeom g;
  if (int rc = make_geom(g, n, st, r0, r1, r2)) return rc;
  ME_REQUIRE(x && c1 && c2 && u, "rigid_apply_chain: null buffer");
  ME_REQUIRE(nbatch >= 1 && nbatch <= 65535, "rigid_apply_chain: bad batch count");
  if (int rc = launch_coef(theta, theta_stride, coef_ws, nbatch, stream)) return rc;
  const void* src[3] = {x, c1, c2};
  const long long sbs[3] = {x_bstride, c_bstride, c_bstride};
  void* dst[3] = {c1, c2, u};
  const long long dbs[3] = {c_bstride, c_bstride, u_bstride};
  for (int l = 0; l < 3; ++l) {
    Bufs b{(const float2*)src[l], sbs[l], 1, (float2*)dst[l], dbs[l], nullptr, 0};
    if (int rc = run_passes(build_apply_chain(l), g, b, nbatch, coef_ws, nullptr, nullptr, false, stream)) return rc;
  }
  return 0;
}

long long me_rigid_dot_partials_size(int nbatch, const int* n, const long long* st) {
  VolGeom g;
  if (nbatch < 1 || make_geom(g, n, st, nullptr, nullptr, nullptr) != 0) return 0;
  long long m = 0;
  for (int a = 0; a < 3; ++a) m = std::max(m, linepass_max_tiles(g, a));
  return m * nbatch;
}

int me_rigid_deriv_apply_dot(const void* x, long long x_bstride, const void* u, long long u_bstride, void* c1,
                             void* c2, long long c_bstride, const void* w, long long w_bstride, void* y,
                             long long y_bstride, void* work, long long work_bstride, int nbatch, const int* n,
                             const long long* st, const double* r0, const double* r1, const double* r2,
                             const double* theta, long long theta_stride, double* coef_ws, int p_idx, int build,
                             double* partials, double* out, int out_stride, hipStream_t stream) {
  VolGeom g;
  if (int rc = make_geom(g, n, st, r0, r1, r2)) return rc;
  ME_REQUIRE(w && partials && out && out_stride >= 1, "rigid_deriv_apply_dot: null w/partials/out");
  ME_REQUIRE(p_idx >= 0 && p_idx <= 5, "rigid_deriv_apply_dot: p_idx must be in [0, 5]");
  ME_REQUIRE(p_idx >= 3 || u, "rigid_deriv_apply_dot: translation columns need u = T x");
  ME_REQUIRE(p_idx < 3 || (work && y), "rigid_deriv_apply_dot: rotation columns need y and work buffers");
  ME_REQUIRE(!build || (x && c1 && c2), "rigid_deriv_apply_dot: building the chain needs x, c1 and c2");
  ME_REQUIRE(p_idx != 3 || x, "rigid_deriv_apply_dot: column 3 needs x");
  ME_REQUIRE(p_idx != 4 || c1, "rigid_deriv_apply_dot: column 4 needs c1");
  ME_REQUIRE(p_idx != 5 || c2, "rigid_deriv_apply_dot: column 5 needs c2");
  ME_REQUIRE(nbatch >= 1 && nbatch <= 65535, "rigid_deriv_apply_dot: bad batch count");
  if (int rc = launch_coef(theta, theta_stride, coef_ws, nbatch, stream)) return rc;
  if (build) {
    Bufs b0{(const float2*)x, x_bstride, 1, (float2*)c1, c_bstride, nullptr, 0};
    if (int rc = run_passes(build_apply_chain(0), g, b0, nbatch, coef_ws, nullptr, nullptr, false, stream)) return rc;
    Bufs b1{(const float2*)c1, c_bstride, 1, (float2*)c2, c_bstride, nullptr, 0};
    if (int rc = run_passes(build_apply_chain(1), g, b1, nbatch, coef_ws, nullptr, nullptr, false, stream)) return rc;
  }
  const void* from = p_idx <= 2 ? u : (p_idx == 3 ? x : (p_idx == 4 ? (const void*)c1 : (const void*)c2));
  const long long fbs = p_idx <= 2 ? u_bstride : (p_idx == 3 ? x_bstride : c_bstride);
  Bufs b{(const float2*)from, fbs, 1, (float2*)y, y_bstride, (float2*)work, work_bstride};
  long long nblk = 0;
  DotOut dot;
  dot.w = (const float2*)w;
  dot.w_bs = w_bstride;
  dot.partial = partials;
  dot.grid_x = &nblk;
  if (int rc = run_passes(build_deriv_apply(p_idx), g, b, nbatch, coef_ws, nullptr, nullptr, false, stream, dot))
    return rc;
  ME_REQUIRE(nblk > 0, "rigid_deriv_apply_dot: no workgroups");
  hipLaunchKernelGGL(k_dot_finalize, dim3(nbatch), dim3(256), 0, stream, partials, nblk, out, out_stride);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_fft_axis(const void* x, long long x_bstride, void* y, long long y_bstride, int nbatch, const int* n,
                const long long* st, int axis, int inverse, hipStream_t stream) {
  VolGeom g;
  if (int rc = make_geom(g, n, st, nullptr, nullptr, nullptr)) return rc;
  ME_REQUIRE(x && y, "fft_axis: null input/output");
  LineArgs A{};
  A.nbatch = nbatch;
  A.mode = inverse ? 2 : 1;
  A.in[0] = (const float2*)x;
  A.in_bs[0] = x_bstride;
  A.out[0] = (float2*)y;
  A.out_bs[0] = y_bstride;
  HostPhase ph[2][2];
  ph[0][0].active = true;
  return launch_linepass(g, axis, 1, 1, A, ph, stream);
}

}  // extern "C"
