// Host-side helpers shared by the C-ABI translation units: error capture and
// the line-pass descriptor used by every transform kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>

namespace me {

// Last error message of the calling host thread (returned by me_last_error).
void set_error(const std::string& msg);
const char* get_error();

#define ME_TRY_HIP(expr)                                                          \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      ::me::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));         \
      return 1;                                                                   \
    }                                                                             \
  } while (0)

#define ME_REQUIRE(cond, msg)          \
  do {                                 \
    if (!(cond)) {                     \
      ::me::set_error(msg);            \
      return 2;                        \
    }                                  \
  } while (0)

// Number of FP64 coefficients per batch row produced from one parameter row
// [q1, q2, q3, t1, t2, t3] (see k_rigid_coef).
constexpr int kNumCoef = 16;
inline int ci_tan(int rho) { return 3 + 4 * rho; }  // tan(theta/2)
inline int ci_sin(int rho) { return 4 + 4 * rho; }  // sin(theta)
inline int ci_cos(int rho) { return 5 + 4 * rho; }  // cos(theta)
inline int ci_g(int rho) { return 6 + 4 * rho; }    // (1 + tan^2(theta/2)) / 2

// Diagonal factor applied between the forward and inverse FFT of one line
// pass, for one (output, input) pair:
//   phi(k, j) = [deriv ? i*(g*r_b[j] + h)*k_a : 1] * exp(i*k_a*(c*r_b[j] + d))
// with c = sc*coef[ic], d = sd*coef[id], g = sg*coef[ig] (absent index -> 0),
// k_a = 2*pi*fftfreq(N)[k] (Nyquist at -pi) and r_b the grid vector of the
// axis b (b = -1: no r dependence).
struct PhaseSpec {
  float sc, sd, sg, h;
  int ic, id, ig;
  int slot;  // device-side: 0 = r_b indexed by the tile coordinate u, 1 = by v, -1 none
  int deriv;
  int active;
};

struct LineArgs {
  const float2* in[2];
  float2* out[2];
  long long in_bs[2], out_bs[2];  // per batch-entry strides (elements)
  long long st_a, st_u, st_v;     // element strides: along the line, tile axis, other axis
  int n_u, n_v;
  int mode;        // 0: FFT, phase, inverse FFT; 1: forward FFT only; 2: inverse FFT only
  int accumulate;  // out += result
  int reduce;      // sum the result over the batch, write batch entry 0 only
  int nbatch;
  const double* coef;  // nbatch x ncoef
  int ncoef;
  const double* r_u;
  const double* r_v;
  const unsigned char* mask;  // optional output mask (same layout as out)
  long long mask_bs;
  // Dot epilogue (single-output MODE 0 passes, not REDUCE): when dot_w is
  // set the output is not stored; each workgroup writes the FP64 partial
  // sum of Re(conj(dot_w) out) over its lines to
  // dot_partial[batch * gridDim.x + blockIdx.x].
  const float2* dot_w;
  long long dot_w_bs;
  double* dot_partial;
  PhaseSpec ph[2][2];  // [output][input]
};

// Host description of a phase, with the b axis as a volume axis index.
struct HostPhase {
  bool active = false;
  int ic = -1;
  float sc = 0.f;
  int id = -1;
  float sd = 0.f;
  int b_axis = -1;
  int deriv = 0;
  int ig = -1;
  float sg = 0.f;
  float h = 0.f;
};

struct VolGeom {
  int n[3];
  long long st[3];
  const double* r[3];  // device grid vectors (may be null if unused)
};

// Launches one line pass along ``axis`` over a batch of volumes.  ``args``
// carries buffers, batch strides, mode/flags and coefficients; geometry and
// phases are filled in here.  Returns 0 or an error code (message set).
int launch_linepass(const VolGeom& g, int axis, int nin, int nout, LineArgs args,
                    const HostPhase ph[2][2], hipStream_t stream, long long* grid_x = nullptr);

// Upper bound of the workgroups per batch entry of a pass along ``axis``
// (the fewest lines per workgroup any variant uses): the size of a dot
// epilogue's partials per batch entry.
long long linepass_max_tiles(const VolGeom& g, int axis);

}  // namespace me
