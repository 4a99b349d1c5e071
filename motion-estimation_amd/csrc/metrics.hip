// Image-quality metrics of the experiment scripts (simulation.py:87-97,
// 476-477) on the GPU: the error volume |corrected - gt| with its 2-norm, and
// the gradient-entropy sharpness proxy (forward differences with a zero last
// difference per axis, gradient magnitude, 256-bin density histogram over
// [0, max + 1e-8], entropy -sum h log2 h with h + 1e-12).  The histogram
// follows NumPy's uniform-bin rule for float32 data: index from
// ((x - lo) / (hi - lo)) * nbins in float32, clamped, then moved by one where
// x falls outside the float32 bin edges it was assigned (edges supplied by
// the caller, computed as NumPy does), so the integer counts match NumPy's.
#include "me_device.h"
#include "me_internal.h"
#include "../../include/motion_estimation_amd.h"

namespace me {

constexpr int kHistBins = 256;
constexpr int kMetricThreads = 256;

// Correctly rounded float32 operations, each evaluated in FP64 and rounded
// once: FP64 carries more than 2*24+2 bits, so add, multiply, divide and
// square root round exactly as IEEE float32 does (NumPy's float32 ufuncs),
// whatever the device's own float32 divide and square-root sequences do.
__device__ __forceinline__ float f32_sub(float a, float b) { return (float)((double)a - (double)b); }
__device__ __forceinline__ float f32_add(float a, float b) { return (float)((double)a + (double)b); }
__device__ __forceinline__ float f32_mul(float a, float b) { return (float)((double)a * (double)b); }
__device__ __forceinline__ float f32_div(float a, float b) { return (float)((double)a / (double)b); }
__device__ __forceinline__ float f32_sqrt(float a) { return (float)sqrt((double)a); }

// NumPy order: sqrt((gx*gx + gy*gy) + gz*gz), differences of neighbours
__device__ __forceinline__ float grad_mag(const float* __restrict__ v, int n0, int n1, int n2, int i0, int i1,
                                          int i2) {
  const long long s0 = (long long)n1 * n2, s1 = n2;
  const long long o = i0 * s0 + i1 * s1 + i2;
  const float c = v[o];
  const float gx = i0 + 1 < n0 ? f32_sub(v[o + s0], c) : 0.f;
  const float gy = i1 + 1 < n1 ? f32_sub(v[o + s1], c) : 0.f;
  const float gz = i2 + 1 < n2 ? f32_sub(v[o + 1], c) : 0.f;
  return f32_sqrt(f32_add(f32_add(f32_mul(gx, gx), f32_mul(gy, gy)), f32_mul(gz, gz)));
}

__global__ void __launch_bounds__(kMetricThreads) k_abs_diff(const float2* __restrict__ a, const float2* __restrict__ b,
                                                             long long n, float* __restrict__ out,
                                                             double* __restrict__ partial) {
  __shared__ double sh[kMetricThreads / 64];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float2 d = make_float2(a[i].x - b[i].x, a[i].y - b[i].y);
    const float m = hypotf(d.x, d.y);
    if (out) out[i] = m;
    acc += (double)m * (double)m;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kMetricThreads / 64; ++w) s += sh[w];
    partial[blockIdx.x] = s;
  }
}

__global__ void k_sum_fixed(const double* __restrict__ partial, int n, double* __restrict__ out) {
  __shared__ double sh[kMetricThreads / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kMetricThreads / 64; ++w) s += sh[w];
    out[0] = s;
  }
}

// max of the gradient magnitude (non-negative floats order as their bits)
__global__ void __launch_bounds__(kMetricThreads) k_grad_max(const float* __restrict__ v, int n0, int n1, int n2,
                                                             unsigned int* __restrict__ out) {
  const long long n = (long long)n0 * n1 * n2;
  float m = 0.f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int i2 = (int)(i % n2), i1 = (int)((i / n2) % n1), i0 = (int)(i / ((long long)n1 * n2));
    m = fmaxf(m, grad_mag(v, n0, n1, n2, i0, i1, i2));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_down(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

__global__ void __launch_bounds__(kMetricThreads) k_grad_hist(const float* __restrict__ v, int n0, int n1, int n2,
                                                              const float* __restrict__ edges,
                                                              unsigned long long* __restrict__ counts) {
  __shared__ unsigned int h[kHistBins];
  __shared__ float e[kHistBins + 1];
  for (int j = threadIdx.x; j < kHistBins; j += blockDim.x) h[j] = 0u;
  for (int j = threadIdx.x; j <= kHistBins; j += blockDim.x) e[j] = edges[j];
  __syncthreads();
  const float lo = e[0], hi = e[kHistBins];
  const float denom = f32_sub(hi, lo);
  const long long n = (long long)n0 * n1 * n2;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int i2 = (int)(i % n2), i1 = (int)((i / n2) % n1), i0 = (int)(i / ((long long)n1 * n2));
    const float x = grad_mag(v, n0, n1, n2, i0, i1, i2);
    if (!(x >= lo && x <= hi)) continue;
    const float f = f32_mul(f32_div(f32_sub(x, lo), denom), (float)kHistBins);
    int k = (int)f;
    if (k == kHistBins) k -= 1;
    if (x < e[k]) k -= 1;
    if (x >= e[k + 1] && k != kHistBins - 1) k += 1;
    atomicAdd(&h[k], 1u);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kHistBins; j += blockDim.x)
    if (h[j]) atomicAdd(&counts[j], (unsigned long long)h[j]);
}

__global__ void k_zero_u64(unsigned long long* __restrict__ c, int n) {
  for (int j = threadIdx.x; j < n; j += blockDim.x) c[j] = 0ull;
}

__global__ void k_zero_u32(unsigned int* __restrict__ c) {
  if (threadIdx.x == 0) c[0] = 0u;
}

// density h_j = count_j / (db_j * total), db_j = e[j+1] - e[j] in float32
// (NumPy: n / np.diff(bin_edges) / n.sum()); entropy -sum (h + 1e-12) log2(h + 1e-12)
__global__ void k_hist_entropy(const unsigned long long* __restrict__ counts, const float* __restrict__ edges,
                               double* __restrict__ out) {
  __shared__ double sh[kHistBins / 64];
  __shared__ double tot_sh;
  const int j = threadIdx.x;
  double c = (double)counts[j];
  double t = c;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
  if ((j & 63) == 0) sh[j >> 6] = t;
  __syncthreads();
  if (j == 0) tot_sh = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  const double db = (double)f32_sub(edges[j + 1], edges[j]);
  const double h = c / db / tot_sh + 1e-12;
  double term = h * log2(h);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) term += __shfl_down(term, o, 64);
  __syncthreads();
  if ((j & 63) == 0) sh[j >> 6] = term;
  __syncthreads();
  if (j == 0) out[0] = -(sh[0] + sh[1] + sh[2] + sh[3]);
}

inline unsigned metric_grid(long long n) {
  long long b = (n + kMetricThreads - 1) / kMetricThreads;
  if (b < 1) b = 1;
  return (unsigned)(b < 4096 ? b : 4096);
}

}  // namespace me

using namespace me;

extern "C" {

int me_metric_partials_size(void) { return 4096; }

int me_abs_diff_norm(const void* a, const void* b, long long n, float* out_abs, double* partials, double* sumsq,
                     hipStream_t stream) {
  ME_REQUIRE(a && b && partials && sumsq && n >= 1, "abs_diff_norm: bad argument");
  const unsigned g = metric_grid(n);
  hipLaunchKernelGGL(k_abs_diff, dim3(g), dim3(kMetricThreads), 0, stream, (const float2*)a, (const float2*)b, n,
                     out_abs, partials);
  hipLaunchKernelGGL(k_sum_fixed, dim3(1), dim3(kMetricThreads), 0, stream, (const double*)partials, (int)g, sumsq);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_gradient_max(const float* vol, const int* n, unsigned int* max_bits, hipStream_t stream) {
  ME_REQUIRE(vol && n && max_bits && n[0] >= 1 && n[1] >= 1 && n[2] >= 1, "gradient_max: bad argument");
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, max_bits);
  hipLaunchKernelGGL(k_grad_max, dim3(metric_grid((long long)n[0] * n[1] * n[2])), dim3(kMetricThreads), 0, stream,
                     vol, n[0], n[1], n[2], max_bits);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

int me_gradient_entropy(const float* vol, const int* n, const float* edges, unsigned long long* counts,
                        double* entropy, hipStream_t stream) {
  ME_REQUIRE(vol && n && edges && counts && entropy && n[0] >= 1 && n[1] >= 1 && n[2] >= 1,
             "gradient_entropy: bad argument");
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(kHistBins), 0, stream, counts, kHistBins);
  hipLaunchKernelGGL(k_grad_hist, dim3(metric_grid((long long)n[0] * n[1] * n[2])), dim3(kMetricThreads), 0, stream,
                     vol, n[0], n[1], n[2], edges, counts);
  hipLaunchKernelGGL(k_hist_entropy, dim3(1), dim3(kHistBins), 0, stream, (const unsigned long long*)counts, edges,
                     entropy);
  ME_TRY_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
