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
#define ME_FLAG_ACCUMULATE 4 /* add the (summed) output to y instead of overwriting it */

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
 * x_parts (1 or 2): input b is the sum of x_parts volumes at
 *   x + b*x_bstride + p*N0*N1*N2 (the coil-group partials of me_encode_adjoint).
 * theta: nbatch rows at theta_stride doubles; coef_ws: device scratch of
 *   16*nbatch doubles.  x and y may alias (in place) unless SUM_BATCH. */
int me_rigid_transform(const void* x, long long x_bstride, int x_parts, void* y, long long y_bstride, void* work,
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

/* The six derivative columns of me_rigid_deriv with the rotation chain shared:
 * c1 = Rz x, c2 = Ry c1, c3 = Tr Rx c2 = T_d x (one volume per batch entry
 * each, at c_bstride) are computed when build != 0 (call p_idx = 0 first with
 * build = 1, then p_idx = 1..5 with build = 0 and the same theta); column
 * p_idx is then produced from the shortest chain level that contains it (a
 * translation column is one -i k_p pass over c3).  The same operator as
 * me_rigid_deriv (transform.py:476-504), equal to FP32 rounding, with 38
 * instead of 66 line passes per six columns. */
int me_rigid_deriv_chain(const void* x, long long x_bstride, void* c1, void* c2, void* c3, long long c_bstride,
                         void* y, long long y_bstride, void* work, long long work_bstride, int nbatch, const int* n,
                         const long long* st, const double* r0, const double* r1, const double* r2,
                         const double* theta, long long theta_stride, double* coef_ws, int p_idx, int build,
                         hipStream_t stream);

/* Jacobian columns of the apply order T = Tr Rz Ry Rx (the composition of
 * me_rigid_transform mode 0, transform.py:267-318) for gradients through
 * RigidTransformCudaOptimzied.apply; the reference's derivative class
 * (me_rigid_deriv) differentiates the reverse order instead.  Columns share
 * the rotation chain c1 = Rx x, c2 = Tr_z Ry c1 (computed when build != 0,
 * batch stride c_bstride) and u = T x (x_bstride/u_bstride as for
 * me_rigid_transform):
 *   p = 0,1,2 from u (one pass), p = 3 from x, p = 4 from c1, p = 5 from c2.
 * work: one volume per entry (rotation columns). */
int me_rigid_deriv_apply(const void* x, long long x_bstride, const void* u, long long u_bstride, void* c1, void* c2,
                         long long c_bstride, void* y, long long y_bstride, void* work, long long work_bstride,
                         int nbatch, const int* n, const long long* st, const double* r0, const double* r1,
                         const double* r2, const double* theta, long long theta_stride, double* coef_ws, int p_idx,
                         int build, hipStream_t stream);

/* Forward transform that keeps the apply-order chain for
 * me_rigid_deriv_apply / _dot with build = 0: c1 = Rx x, c2 = Tr_z Ry c1 and
 * u = T x = Tr_y (Rz Tr_x) c2, the passes of me_rigid_transform mode 0
 * (transform.py:263-318) in its order, so u equals its output bit for bit. */
int me_rigid_apply_chain(const void* x, long long x_bstride, void* c1, void* c2, long long c_bstride, void* u,
                         long long u_bstride, int nbatch, const int* n, const long long* st, const double* r0,
                         const double* r1, const double* r2, const double* theta, long long theta_stride,
                         double* coef_ws, hipStream_t stream);

/* The same columns reduced against w without storing them:
 *   out[b * out_stride] = Re <w_b, column p_idx of entry b>   (FP64)
 * The last line pass of the column sums Re(conj(w) col) per workgroup into
 * `partials` (me_rigid_dot_partials_size doubles) and a fixed-order sum per
 * entry follows.  y (one volume per entry) holds the rotation columns'
 * intermediate passes (unused for p_idx < 3).  This is the motion gradient
 * of a loss through RigidTransformCudaOptimzied.apply when w is the adjoint
 * encoding of the residual; it corresponds to the reference's gradient
 * entries g_p = Re <r, J_p> (transform_estimation.py:56), formed in image
 * space. */
int me_rigid_deriv_apply_dot(const void* x, long long x_bstride, const void* u, long long u_bstride, void* c1,
                             void* c2, long long c_bstride, const void* w, long long w_bstride, void* y,
                             long long y_bstride, void* work, long long work_bstride, int nbatch, const int* n,
                             const long long* st, const double* r0, const double* r1, const double* r2,
                             const double* theta, long long theta_stride, double* coef_ws, int p_idx, int build,
                             double* partials, double* out, int out_stride, hipStream_t stream);

/* Doubles of the `partials` workspace of me_rigid_deriv_apply_dot. */
long long me_rigid_dot_partials_size(int nbatch, const int* n, const long long* st);

/* Orthonormal 1-D FFT (inverse != 0: inverse FFT) along `axis` of a batch of
 * volumes; x and y may alias.  Replaces the cupyx.scipy.fft fft/ifft/fftn
 * calls used to prepare data (joint_recon.py:63-68, pyramid.py:65). */
int me_fft_axis(const void* x, long long x_bstride, void* y, long long y_bstride, int nbatch, const int* n,
                const long long* st, int axis, int inverse, hipStream_t stream);


/* ---------------------------------------------------------------- tuning
 * Lines per workgroup of the plain line pass (one input, one output, FFT ->
 * phase -> inverse FFT; the bulk of the rigid-transform passes): 8, 16
 * (default) or 32, kept in a process-wide table keyed by the line geometry
 * (line length, strides, tile extents) of a volume with extents n and element
 * strides st along `axis`.  No counterpart in the reference (CuPy picks its
 * own FFT plans); results are bitwise independent of the choice. */

/* Set the lines per workgroup for passes along `axis` of this geometry
 * (0 restores the default).  Errors if that count is not built for the line
 * length (8 and 32 exist up to 512 threads per workgroup). */
int me_line_tuning_set(const int* n, const long long* st, int axis, int lines);

/* Lines per workgroup in use for passes along `axis` of this geometry. */
int me_line_tuning_get(const int* n, const long long* st, int axis);

/* Forget every tuned entry. */
void me_line_tuning_clear(void);

/* Autotuner: for each axis, time `reps` plain passes x -> y (nbatch volumes
 * at bstride elements; y may equal x, its contents are overwritten) with 8,
 * 16 and 32 lines per workgroup using HIP events on `stream`, and keep the
 * fastest in the table.  chosen[3] (nullable) receives the picks,
 * times_ms[9] (nullable) the mean time per pass of each candidate (-1 where
 * not built).  Synchronises the stream; not for use under graph capture. */
int me_tune_line_passes(const void* x, void* y, long long bstride, int nbatch, const int* n, const long long* st,
                        int reps, int* chosen, float* times_ms, hipStream_t stream);

/* ---------------------------------------------------------------- graphs
 * Thin wrappers over HIP stream capture for replaying a recorded sequence of
 * the calls above (no counterpart in the reference, which launches eagerly
 * from the host).  Between me_graph_begin and me_graph_end every call on the
 * stream (and on streams joined to it through events) is recorded, not run;
 * allocations on the capturing thread are errors. */
int me_graph_begin(hipStream_t stream);
/* Ends the capture and instantiates it: *exec receives the graph handle. */
int me_graph_end(hipStream_t stream, void** exec);
/* Ends a capture left open by an error, discarding it. */
int me_graph_abort(hipStream_t stream);
int me_graph_launch(void* exec, hipStream_t stream);
int me_graph_destroy(void* exec);

/* ---------------------------------------------------------------- encoding
 * Sample plan of a set of shots (host-built from the shot masks
 * (S, N0, N1, 1), simulation.py:396-402): samples (k0, k1) of every shot in one
 * global list, each shot's run contiguous and sorted by (k1, k0);
 * shot_off/shot_cnt give each shot's run; colptr holds, per shot, N1+1
 * column pointers relative to its run.  Gathered hybrid data of a shot live
 * at Y + shot_off*N2*C, laid out [x2][coil][sample].
 * Images and maps are in plane-major layout (see above); maps are C volumes. */

/* Size (doubles) of the per-workgroup partials workspace of me_encode_forward
 * for one shot; multiply by nbatch. */
long long me_encode_partials_size(int n2, int ncoils, int max_cnt, int n1);

/* Gathered forward encoding of a batch of shots:
 *   Y_s = M_s F_01 (S_c u_s)      (transform_estimation.py:42-45 without the
 *   readout FFT, which is applied once to the measured data)
 * u: image per shot at u_bstride (0: one image for all shots).  If y_meas is
 * given, Y - y_meas is produced (the residual) unless flags bit 0
 * (ME_ENC_STORE_RAW) asks for Y itself.  y (nullable) receives the values;
 * norms (nullable, needs partials) receives ||Y - y_meas||^2 per shot in FP64
 * (the objective, joint_recon.py:108-115, when y is the prediction).
 * max_cnt: largest shot sample count in the batch. */
#define ME_ENC_STORE_RAW 1
int me_encode_forward(const void* u, long long u_bstride, const void* maps, int ncoils, int nbatch, const int* n,
                      const int* samp_k0, const int* samp_k1, const int* shot_off, const int* shot_cnt, int max_cnt,
                      void* y, const void* y_meas, double* partials, double* norms, int flags, hipStream_t stream);

/* Allocation-free variant of me_encode_forward.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as me_encode_forward without `partials`: the
 * per-workgroup FP64 norm partials live in a library-owned static device
 * array of ME_ENCODE_STATIC_PARTIALS doubles (8 MiB, one per GPU, held until
 * the process ends).  Needs nbatch * me_encode_partials_size(N2, ncoils,
 * max_cnt, N1) <= ME_ENCODE_STATIC_PARTIALS when norms is given, else it
 * fails with an error before any device work.  `stream` must belong to the
 * current device.  Every call on one GPU shares that GPU's array: calls that
 * overlap (concurrent streams) race.  Calls no hipMalloc or hipFree. */
#define ME_ENCODE_STATIC_PARTIALS (1 << 20)
int me_encode_forward_static(const void* u, long long u_bstride, const void* maps, int ncoils, int nbatch,
                             const int* n, const int* samp_k0, const int* samp_k1, const int* shot_off,
                             const int* shot_cnt, int max_cnt, void* y, const void* y_meas, double* norms, int flags,
                             hipStream_t stream);

/* Forward-encoding variants and their tuning (no counterpart in the
 * reference).  k_encode_fwd runs G coils per workgroup with SPT samples per
 * thread: variant 0 = 16 x 2 (default), 1 = 8 x 2, 2 = 8 x 4, 3 = 4 x 4; fewer
 * coils per group waste fewer threads when the coil count is not a multiple
 * of 16 but split large shots into more sample chunks, each of which redoes
 * the row FFTs.  The choice is kept in a process-wide table keyed by (N0, N1, N2,
 * ncoils, max_cnt).  The samples are bitwise the same for every variant; the
 * FP64 norms agree to rounding (their per-workgroup partial sums group the
 * samples differently).  me_encode_partials_size covers every variant. */
int me_encode_tuning_set(const int* n, int ncoils, int max_cnt, int variant);
/* Variant in use for this geometry (0 unless tuned). */
int me_encode_tuning_get(const int* n, int ncoils, int max_cnt);
void me_encode_tuning_clear(void);
/* Times `reps` launches of me_encode_forward (no residual, no norms; y is
 * overwritten) for each variant with HIP events on `stream`, keeps the
 * fastest in the table; chosen (nullable) receives it, times_ms[4]
 * (nullable) the mean time per launch of each.  Synchronises the stream;
 * not for use under graph capture. */
int me_tune_encode_forward(const void* u, long long u_bstride, const void* maps, int ncoils, int nbatch, const int* n,
                           const int* samp_k0, const int* samp_k1, const int* shot_off, const int* shot_cnt,
                           int max_cnt, void* y, int reps, int* chosen, float* times_ms, hipStream_t stream);

/* Number of partial images per shot written by me_encode_adjoint (1: the
 * adjoint sums all coils itself). */
int me_encode_coil_groups(int ncoils);

/* Adjoint of me_encode_forward for a batch of shots:
 *   out_s = sum_c conj(S_c) F_01^H (M_s Y_{s,c})   (image_estimation.py:94-102,
 *   before T^H), written at out + s*out_bstride (plane-major). */
int me_encode_adjoint(const void* y, const void* maps, int ncoils, int nbatch, const int* n, const int* samp_k0,
                      const int* shot_off, const int* shot_cnt, const int* colptr, void* out, long long out_bstride,
                      hipStream_t stream);

/* y[c] = scale * x[(c + offset) mod x_dims] over 4-D index spaces (coil, x0,
 * x1, x2) with arbitrary strides, y_dims <= x_dims.  Used for the reference's
 * ifftshift/fftshift (joint_recon.py:63-65, 106), SigPy-style centred crops
 * (pyramid.py:65-67) and layout changes. */
int me_permute_shift(const void* x, const long long* x_strides, const int* x_dims, void* y,
                     const long long* y_strides, const int* y_dims, const int* offset, float scale,
                     hipStream_t stream);

/* P = 1/(sum_c |S_c|^2 + eps), M = sum_c |S_c|^2 > thr (joint_recon.py:83-85);
 * either output may be null. */
int me_coil_ssq(const void* maps, int ncoils, long long vol, float eps, float thr, float* P, unsigned char* M,
                hipStream_t stream);

/* Gather measured hybrid data: Ymeas[shot_off*N2*C + (x2*C + c)*cnt + w] =
 * scale * ksp_h[c][k0_w][k1_w][x2], ksp_h in the reference C order after the
 * readout inverse FFT. */
int me_gather_hybrid(const void* ksp_h, int ncoils, const int* n, const int* samp_k0, const int* samp_k1,
                     const int* shot_off, const int* shot_cnt, int nshots, long long total_samples, float scale,
                     void* ymeas, hipStream_t stream);

/* ksp_h[c][k0_w][k1_w][x2] += Y[...]: scatter of gathered hybrid data back to
 * dense C-order hybrid k-space (atomic adds).  Used to synthesise the
 * motion-corrupted k-space, sum_s M_s F S T_s x (simulation.py:65-85). */
int me_scatter_hybrid(const void* y, int ncoils, const int* n, const int* samp_k0, const int* samp_k1,
                      const int* shot_off, const int* shot_cnt, int nshots, long long total_samples, void* ksp_h,
                      hipStream_t stream);

/* *out_bits = max(*out_bits, bits of max |x_i|) (atomic; zero it first). */
int me_max_abs(const void* x, long long n, unsigned int* out_bits, hipStream_t stream);

/* ---------------------------------------------------------------- CG / LM
 * CG state: double[10] = {rsold, rsnew, pAp, alpha, beta, done, iters, rule,
 * tol0, -}; the caller sets rule (and tol0) before me_cg_init:
 *   rule 0: estimate_image_cg (image_estimation.py:156-157): stop after an
 *           update when sqrt|<r, P r>| < tol;
 *   rule 1: sigpy.alg.ConjugateGradient (sigpy.mri.app.SenseRecon,
 *           simulation.py:462-466): stop before the first update when
 *           sqrt(<r, P r>) <= tol0, on <p, A p> <= 0 without updating, and
 *           after an update when sqrt(<r, P r>) <= tol. */

/* Number of doubles in the partials workspace of the vector kernels. */
int me_vec_partials_size(void);

/* *out = Re sum_i conj(a_i) b_i (w_i if w) in FP64 (cupy.vdot(...).real). */
int me_dot_re(const void* a, const void* b, const float* w, long long n, double* partials, double* out,
              hipStream_t stream);

/* r = b - Ap; p = P r; rsold = Re<r, P r>; done = 0 (image_estimation.py:133-138). */
int me_cg_init(const void* b, const void* Ap, void* r, void* p, const float* P, long long n, double* state,
               double* partials, hipStream_t stream);

/* One PCG update after Ap = A p (image_estimation.py:143-163): pAp, alpha,
 * x += alpha p, r -= alpha Ap, rsnew = Re<r, P r>, the tolerance test, beta,
 * p = P r + beta p.  No-op once the tolerance test has fired.  Optional
 * tracking (yx, yp non-null, ny values): yx += alpha * yp with the same alpha,
 * so that yx = E T x stays current when yp = E T p (gathered encodings). */
int me_cg_step(void* x, void* r, void* p, const void* Ap, const float* P, long long n, double tol, double* state,
               double* partials, void* yx, const void* yp, long long ny, hipStream_t stream);

/* Per-entry FP64 inner products out[b * out_stride] = Re <a_b, c_b> =
 * sum_i Re(conj(a_b[i]) c_b[i]) over n complex values of each of nbatch
 * entries (deterministic fixed-order sums).  partials: at least
 * me_dot_batched_partials_size(nbatch) doubles.  Used for the per-shot
 * parameter gradients (no counterpart in the reference, which has no
 * gradient path). */
long long me_dot_batched_partials_size(int nbatch);
int me_dot_re_batched(const void* a, long long a_bstride, const void* c, long long c_bstride, long long n, int nbatch,
                      double* partials, double* out, long long out_stride, hipStream_t stream);

/* out (+)= sum over nbatch entries (x + b*x_bstride) of n complex values,
 * in a fixed order (accumulate != 0 adds to out). */
int me_sum_batch(const void* x, long long x_bstride, int nbatch, long long n, void* out, int accumulate,
                 hipStream_t stream);

/* y += a x for complex vectors and a real scalar. */
int me_axpy_real(void* y, double a, const void* x, long long n, hipStream_t stream);

/* out = a - b (complex, n values). */
int me_csub(void* out, const void* a, const void* b, long long n, hipStream_t stream);

/* x *= mask (uint8) (joint_recon.py:100). */
int me_mask_mul(void* x, const unsigned char* mask, long long n, hipStream_t stream);

/* Doubles of LM gram partials for nbatch shots. */
long long me_lm_partials_size(int nbatch);

/* Per shot: stats[28] = {||r||^2, Re J_p^H r (6), Re J_i^H J_j (i<=j, 21)} over
 * its gathered samples in FP64 (transform_estimation.py:47-68).  r and the
 * columns use the gathered layout; column p at J + p*jstride; per_sample =
 * N2*C values per sample. */
int me_lm_gram(const void* r, const void* J, long long jstride, const int* shot_off, const int* shot_cnt,
               int per_sample, int nbatch, double* partials, double* stats, hipStream_t stream);

/* (Gram + damp I) delta = g, cand = theta - delta (transform_estimation.py:71, 125). */
int me_lm_solve(const double* stats, const double* damp, const double* theta, int theta_stride, double* cand,
                double* delta, int nbatch, hipStream_t stream);

/* Geodesic acceleration, the reference's experimental option
 * (transform_estimation.py:73-121).  me_lm_ga_accumulate: per shot b,
 * out_b = (accumulate ? out_b : 0) + coef * v_b[p_idx] * d_b over n values,
 * v = -delta; called with the derivative image of column p at theta + h v
 * (coef = 1/(2h)) and at theta - h v (coef = -1/(2h)) for every p, it builds
 * the image whose encoding is J'v, the directional derivative of the Jacobian
 * (transform_estimation.py:78-100 estimates it by a forward difference at
 * eps = 1e-4 of FP32 columns; the central difference at h = 1e-3 estimates
 * the same quantity with a tenth of the rounding noise).
 * me_lm_ga_solve: with stats_dir = me_lm_gram of J'v
 * against J, rhs_p = -Re<J'v, J_p>, a = (Gram + damp I)^-1 rhs, and
 * cand = theta + v + a/2 when |a| <= |v|, else theta + v
 * (transform_estimation.py:102-121). */
int me_lm_ga_accumulate(const void* d, long long d_bstride, long long n, int nbatch, const double* delta, int p_idx,
                        double coef, int accumulate, void* out, long long out_bstride, hipStream_t stream);
int me_lm_ga_solve(const double* stats, const double* stats_dir, const double* damp, const double* theta,
                   int theta_stride, const double* delta, double* cand, int nbatch, hipStream_t stream);

/* Accept iff e1 < e0: theta = cand, damp = max(damp/5, 1e-4), flags set if
 * |theta - cand| < tol; else damp = min(1.5 damp, 1e16)
 * (transform_estimation.py:138-144).  accepted (nullable) records decisions. */
int me_lm_accept(const double* stats, const double* e1, double* theta, int theta_stride, const double* cand,
                 double* damp, unsigned char* flags, unsigned char* accepted, double tol, int nbatch,
                 hipStream_t stream);

/* mean[6] = mean over shots of theta; theta -= mean (transform_estimation.py:147-150). */
int me_lm_mean(double* theta, int theta_stride, int nshots, double* mean, hipStream_t stream);

/* ---------------------------------------------------------------- host
 * DISORDER order (simulation.py:201-245), bit-exact with the reference's
 * stdlib random.Random(seed).shuffle.  Host function, no device work. */
int me_disorder_order(int max_line_number, int max_partition_number, int pat_lines_to_measure, int block_lin,
                      int block_par, unsigned long long seed, int* out_line, int* out_part, long long capacity,
                      long long* out_count, int* out_L, int* out_P);

/* ---------------------------------------------------------------- calibration
 * Preprocessing before the hot path: noise whitening (noise_whiten.py:5-14)
 * and coil compression (coil_compression.py:6-46).  FP64 complex matrices are
 * interleaved (re, im) doubles, row-major, ncoils <= 64. */

/* Doubles of the partials workspace of me_coil_gram. */
long long me_coil_gram_partials_size(int ncoils);

/* gram[i][j] = scale * sum_v w(v) x_i(v) conj(x_j(v)) over n values per coil
 * (x: ncoils x n complex64, coil-major; w nullable real weights).  The noise
 * covariance of noise_whiten.py:7 (w = null, scale = 1/n) and the matrix P of
 * coil_compression.py:17-27 (w = 1/(sum |S|^2 + reg), scale = 1). */
int me_coil_gram(const void* x, int ncoils, long long n, const float* w, double scale, double* partials,
                 double* gram, hipStream_t stream);

/* W = (L^-1)^H with L the lower Cholesky factor of gram + reg I
 * (noise_whiten.py:8-10, as the reference forms it). */
int me_whitening_matrix(const double* gram, int ncoils, double reg, double* W, hipStream_t stream);

/* Hermitian eigendecomposition A = V diag(eigvals) V^H (cyclic complex
 * Jacobi, FP64), eigenvalues descending, eigenvectors as the columns of
 * eigvecs (n x n).  For the Hermitian positive semidefinite P of
 * coil_compression.py:30 this is its SVD (U = V, S = eigvals); eigenvector
 * phases are those of the rotations, not LAPACK's. */
int me_hermitian_eig(const double* A, int n, int max_sweeps, double* eigvals, double* eigvecs, hipStream_t stream);

/* y[o] = sum_c M[o][c] x[c] for every voxel (M: cout x cin complex64,
 * row-major; x: cin x n, y: cout x n).  Applies the whitening matrix
 * (noise_whiten.py:11-12) and the compression matrix (coil_compression.py:38-43). */
int me_coil_mix(const void* M, int cout, int cin, const void* x, long long n, void* y, hipStream_t stream);

/* ---------------------------------------------------------------- ESPIRiT
 * rx_espirit_3d (espirit.py:47-239) in four steps: calibration matrix,
 * its Gram matrix, Jacobi eigendecomposition of the Gram matrix (the SVD of
 * espirit.py:114), then the maps from the projector onto the kept subspace.
 * calib_dims = {NKx, NKy, NKz, NRx}, kernel = {kx, ky, kz}, imsize = {Nx, Ny, Nz}. */

/* Complex64 elements of the calibration matrix (kx ky kz NRx rows, one column
 * per kernel-sized block); 0 when the kernel does not fit. */
long long me_espirit_hankel_size(const int* calib_dims, const int* kernel);

/* H (row-major) from calib (NKx, NKy, NKz, NRx) complex64: row k + K c,
 * k = ix + kx (iy + ky iz), columns ordered x block start fastest
 * (espirit.py:9-45: hankel_fwd then fold_rx). */
int me_espirit_calibration_matrix(const void* calib, const int* calib_dims, const int* kernel, void* H,
                                  hipStream_t stream);

/* G[i][j] = sum_b X[i][b] conj(X[j][b]) for X m x n complex64 row-major,
 * G m x m interleaved FP64. */
int me_gram_rows(const void* X, int m, long long n, double* G, hipStream_t stream);

/* Doubles of the workspace of me_jacobi_eig. */
long long me_jacobi_workspace_size(int m);

/* In-place parallel two-sided Jacobi on the Hermitian G (m x m interleaved
 * FP64, overwritten by its diagonalised form): eigvals[i] = eigenvalue i
 * (unsorted), column i of V (m x m) its eigenvector.  Sweeps until the
 * off-diagonal Frobenius norm is <= tol times the diagonal one or
 * max_sweeps; *sweeps_done (nullable) receives the count.  Synchronises
 * the stream once per sweep; work: me_jacobi_workspace_size(m) doubles. */
int me_jacobi_eig(double* G, double* V, int m, int max_sweeps, double tol, double* eigvals, double* work,
                  int* sweeps_done, hipStream_t stream);

/* Largest m for which me_jacobi_eig keeps the rows of V in LDS and applies 32
 * rounds of rotations per pass (default and maximum 8192); above it V is
 * rotated in global memory inside every round.  Returns the previous value.
 * Process-wide; no counterpart in the reference (a control for exercising
 * the large-matrix path at small sizes). */
int me_jacobi_set_lds_rows(int rows);

/* P = V_r V_r^H for the r columns idx[0..r) (device ints) of V (m rows of
 * ldv interleaved FP64 complex): the projector onto the kept singular
 * subspace of espirit.py:118-137.  Xwork holds m x r complex64. */
int me_espirit_projector(const double* V, int m, int ldv, const int* idx, int r, void* Xwork, double* P,
                         hipStream_t stream);

/* C (m x n) = op(A) op(B) in FP64 complex (interleaved, row-major, contiguous):
 * op 0 uses the matrix as stored (A m x k, B k x n), op 1 its conjugate
 * transpose (A stored k x m, B stored n x k).  Subspace iteration for
 * ESPIRiT's kept subspace. */
int me_zgemm(int opa, int opb, int m, int n, int k, const double* A, const double* B, double* C, hipStream_t stream);

/* n complex FP64 standard normal values (real and imaginary parts N(0,1))
 * from a counter-based hash of (seed, index): reproducible start blocks. */
int me_random_normal(double* x, long long n, unsigned long long seed, hipStream_t stream);

/* Column j of X (rows x cols FP64 complex, row-major) times s[j]. */
int me_scale_columns(double* X, long long rows, int cols, const double* s, hipStream_t stream);

/* Complex64 elements of the workspace of me_espirit_maps. */
long long me_espirit_maps_workspace_size(int NRx, const int* kernel, const int* imsize);

/* Maps (Nx, Ny, Nz, NRx) complex64 from the projector P (m = kx ky kz NRx):
 * per voxel the principal eigenvector of the coil matrix of the zero-padded
 * centred inverse FFT of the kept kernels (espirit.py:146-198), zero where
 * sqrt(lambda_max) * prod(imsize) / sqrt(prod(kernel)) <= mask_thresh, with
 * the phase of coil 0 removed (espirit.py:200-239).  Power iteration of at
 * most iters steps.  NRx <= 64. */
int me_espirit_maps(const double* P, int NRx, const int* kernel, const int* imsize, double mask_thresh, int iters,
                    void* work, void* sens, hipStream_t stream);

/* ---------------------------------------------------------------- metrics
 * Image-quality figures of the experiment scripts (simulation.py:87-97,
 * 476-477). */

/* Doubles of the partials workspace of me_abs_diff_norm. */
int me_metric_partials_size(void);

/* out_abs[i] = |a_i - b_i| (complex64 in, float32 out; nullable) and
 * *sumsq = sum of its squares in FP64: err_volume and err_img^2 of
 * simulation.py:476-477. */
int me_abs_diff_norm(const void* a, const void* b, long long n, float* out_abs, double* partials, double* sumsq,
                     hipStream_t stream);

/* *max_bits = bit pattern of the largest gradient magnitude of vol (float32,
 * C order, extents n[3]): sqrt(gx^2 + gy^2 + gz^2) with forward differences
 * and a zero last difference per axis (simulation.py:91-94). */
int me_gradient_max(const float* vol, const int* n, unsigned int* max_bits, hipStream_t stream);

/* Gradient entropy of simulation.py:87-97: 256-bin histogram of the gradient
 * magnitude over the float32 bin edges edges[257] (NumPy's linspace over
 * [0, max + 1e-8]), binned by NumPy's uniform-bin rule (counts[256] receive
 * the integer counts), density over float32 bin widths, and
 * *entropy = -sum (h + 1e-12) log2(h + 1e-12) in FP64. */
int me_gradient_entropy(const float* vol, const int* n, const float* edges, unsigned long long* counts,
                        double* entropy, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* MOTION_ESTIMATION_AMD_H */
