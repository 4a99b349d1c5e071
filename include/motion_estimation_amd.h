/* C ABI of the MI355X (gfx950) joint motion/image reconstruction library
 * (libmeamd.so).  Plain pointers and sizes only.
 *
 * Conventions
 *  - Every compute entry point takes device pointers and an explicit
 *    hipStream_t, enqueues work on that stream without host synchronisation,
 *    allocates nothing, and returns 0 on success or a non-zero code whose
 *    message is returned by me_last_error() (per host thread).
 *  - Complex data is interleaved complex64 (float re, im).
 *  - A volume has extents n[3] = {N0, N1, N2} (reference axes x, y, z) and
 *    element strides st[3]; the reference's C-order layout is
 *    st = {N1*N2, N2, 1}.  The library's plane-major layout, used inside the
 *    solvers, is st = {N1, 1, N0*N1}: each readout (axis-2) plane is a
 *    contiguous N0 x N1 block.
 *  - A "batch" is a set of volumes (one per shot) at a fixed element stride.
 *  - Grid vectors r0/r1/r2 (device, FP64, length N0/N1/N2) are the 1-D factors
 *    of compute_transform_grids_voxel (utils.py:114-116): r in full-resolution
 *    voxel units.  k is 2*pi*fftfreq(N) with the Nyquist bin at -pi
 *    (utils.py:118-120), implied by the extent; only even extents from
 *    {8,12,16,24,32,48,64,96,128,192,256,384,512} are supported.
 *  - Motion parameters are FP64 rows [q1, q2, q3, t1, t2, t3] (translations
 *    in voxels of the level grid, rotations in radians), as the reference's
 *    `transforms` array (joint_recon.py:70-73).
 */
#ifndef MOTION_ESTIMATION_AMD_H
#define MOTION_ESTIMATION_AMD_H

#include <hip/hip_runtime.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ME_ABI_VERSION 1

#define ME_FLAG_SUM_BATCH 1 /* adjoint: sum the batch into one output volume */
#define ME_FLAG_MASK 2      /* multiply the (summed) output by a uint8 mask */

/* Message of the last failed call on this host thread. */
const char* me_last_error(void);
int me_version(void);

/* Rigid transform of a batch of volumes.
 * Replaces RigidTransformCudaOptimzied.apply / .adjoint (transform.py:263, 320).
 *   mode 0: y_b = T(theta_b) x_b            (apply, transform.py:263-318)
 *   mode 1: y_b = T(theta_b)^H x_b          (adjoint, transform.py:320-377)
 *   mode 2: apply with the translation conjugated (inverse=True, transform.py:308)
 * flags ME_FLAG_SUM_BATCH (mode 1): y = sum_b T_b^H x_b, batch intermediates in
 *   `work` (work_bstride); with ME_FLAG_MASK, y is multiplied by mask (uint8,
 *   same layout as y).  This is the shot sum of image_estimation.py:128-129.
 * theta: nbatch rows at theta_stride doubles; coef_ws: device scratch of
 *   16*nbatch doubles.  x and y may alias (in place) unless SUM_BATCH. */
int me_rigid_transform(const void* x, long long x_bstride, void* y, long long y_bstride, void* work,
                       long long work_bstride, int nbatch, const int* n, const long long* st, const double* r0,
                       const double* r1, const double* r2, const double* theta, long long theta_stride,
                       double* coef_ws, int mode, int flags, const unsigned char* mask, hipStream_t stream);

/* Derivative of T_d = Tr Rx Ry Rz with respect to parameter p_idx in [0, 5].
 * Replaces RigidTransformDerivativeCuda.apply (transform.py:476-504), rotation
 * order z, y, x as the reference.  `work` (one volume per batch entry) is
 * needed for p_idx >= 3.  Out-of-range p_idx is an error here; the Python
 * mirror reproduces the reference's copy-through (transform.py:476-504). */
int me_rigid_deriv(const void* x, long long x_bstride, void* y, long long y_bstride, void* work,
                   long long work_bstride, int nbatch, const int* n, const long long* st, const double* r0,
                   const double* r1, const double* r2, const double* theta, long long theta_stride,
                   double* coef_ws, int p_idx, hipStream_t stream);

/* Orthonormal 1-D FFT (inverse != 0: inverse FFT) along `axis` of a batch of
 * volumes; x and y may alias.  Replaces the cupyx.scipy.fft fft/ifft/fftn
 * calls used to prepare data (joint_recon.py:63-68, pyramid.py:65). */
int me_fft_axis(const void* x, long long x_bstride, void* y, long long y_bstride, int nbatch, const int* n,
                const long long* st, int axis, int inverse, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* MOTION_ESTIMATION_AMD_H */
