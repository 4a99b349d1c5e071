"""CPU oracle for the joint motion/image reconstruction hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports, links or
executes this module.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only to check results (or, for
``cpu_baseline``, as the timed CPU restatement), never as the thing measured on
the GPU.

It is a plain NumPy/SciPy restatement of the reference's algorithm, following
the reference op by op (dense masks, 3-D FFTs, dense inner products), in one of
two precisions:

* ``prec="f64"``: complex128 everywhere (the FP64 CPU oracle);
* ``prec="c64"``: the reference's own dtypes (complex64 data, complex64 shear
  phases in ``apply``, complex128 phases rounded at the multiply elsewhere,
  FP32 CG dot products, FP64 6x6 LM quantities).

Each function cites the reference file:line it restates.  The reference itself
cannot run here (CuPy/SigPy are absent and its CG/LM path is CUDA-only); every
function is pinned by the known-answer tests, invariants and finite-difference
checks in ``tests/test_oracle.py``.  Parity status per function:

==============================  ===============================================
function                        pinned by
==============================  ===============================================
compute_transform_grids_voxel   closed form (fftfreq) test
rigid_apply / rigid_adjoint     integer roll KAT, 90-degree permutation KAT,
                                unitarity, adjoint identity, Gaussian rotation
                                vs analytic, rotation-order test
rigid_deriv                     central finite differences of the reverse-order
                                composition Tr*Rx*Ry*Rz, out-of-range p_idx
rigid_deriv_apply               central finite differences of rigid_apply
dc_loss / dc_loss_grad          central finite differences of dc_loss along
                                random image and parameter directions
normal operator / CG            explicit matrix from basis vectors + textbook
                                CG; closed form sum_s E_s^H E_s = I
estimate_transform              6x6 solve vs numpy.linalg.solve on explicit
                                Jacobian; translation-only convergence KAT;
                                Jacobian vs finite differences of the residual
ga_direction / ga_step          closed form of J'v for translation directions
                                (-(k.v)^2 in Fourier); gate and half-step by
                                hand on a constructed system
make_disorder_order             orders written by the reference's own
                                make_disorder_order for 7 tile settings
                                (tests/golden/disorder_orders.npz), bit-exact;
                                tile/union invariants; first tile equals a
                                direct stdlib ``random.shuffle``
disorder_masks                  hand-built masks including the dropped
                                remainder of the order
pyramid resampling              SigPy centred fft/resize restated; crop KATs
noise_whiten                    inv(W^H) is lower triangular with a positive
                                diagonal and squares to cov + reg I
coil_compression                rank-2 maps: r from the index rule (1 coil at
                                perc above the first share), A spans the
                                dominant eigenvector of P
rx_espirit_3d                   Hankel structure of a single nonzero sample;
                                maps from calibration data of known smooth
                                maps agree with them up to the coil-0 phase;
                                unit norm and real coil 0 inside the mask
sense_recon                     full-sampling closed form (A^H A = I: one
                                step); explicit-matrix solve; zero data
                                stops before the first step
gradient_entropy                constant volume and single-axis ramp: counts
                                and entropy in closed form
error_volume                    hand-computed |a - b| and norm
shot_masks_from_counters        hand-built acquisition order, remainder drop
generate_corrupted_kspace       zero motion: masked brute-force centred DFT of
                                S*gt; integer per-shot translations: each
                                shot's circular shift on its own samples
JointRecon                      identity operator (one CG step returns gt,
                                max_norm = max|S gt|, objective 0); single-coil
                                zero-filled projection (ksp/mask shifts);
                                objective of data in the null space of each
                                E_s^H = ||d||^2 / max_norm^2; P and M values of
                                a piecewise coil sum; x zero off M after the
                                mean re-registration
pyramid_reconstruction          per-level JointRecon replaced by a recorder:
                                level shapes, grids, iteration counts, t0 of
                                each level (carried unscaled), final CG with
                                the last level's motion returning gt; constant
                                maps resampled to c*sqrt(N/n), DC sample and
                                mask crops
==============================  ===============================================
"""
from __future__ import annotations

import math
import random
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

try:  # scipy.fft is faster and multithreaded; numpy.fft is the fallback
    import scipy.fft as _sfft

    def _fft(a, axis, workers):
        return _sfft.fft(a, axis=axis, norm="ortho", workers=workers)

    def _ifft(a, axis, workers):
        return _sfft.ifft(a, axis=axis, norm="ortho", workers=workers)

    def _fftn(a, axes, workers):
        return _sfft.fftn(a, axes=axes, norm="ortho", workers=workers)

    def _ifftn(a, axes, workers):
        return _sfft.ifftn(a, axes=axes, norm="ortho", workers=workers)

except ImportError:  # pragma: no cover
    def _fft(a, axis, workers):
        return np.fft.fft(a, axis=axis, norm="ortho")

    def _ifft(a, axis, workers):
        return np.fft.ifft(a, axis=axis, norm="ortho")

    def _fftn(a, axes, workers):
        return np.fft.fftn(a, axes=axes, norm="ortho")

    def _ifftn(a, axes, workers):
        return np.fft.ifftn(a, axes=axes, norm="ortho")


WORKERS = 1


def set_workers(n: int) -> None:
    """Number of threads scipy.fft may use (the CPU baseline sets this)."""
    global WORKERS
    WORKERS = max(1, int(n))


def _cdt(prec):
    return np.complex128 if prec == "f64" else np.complex64


def _rdt(prec):
    return np.float64 if prec == "f64" else np.float32


# ---------------------------------------------------------------------------
# a1: grids  (utils.py:107-137)
# ---------------------------------------------------------------------------
def compute_transform_grids_voxel(shape, voxel_size, downsample_res=None):
    """utils.py:107-137.  r in full-resolution voxel units (fftshifted
    linspace over the full extent, sampled at the level size); k the
    fftshifted linspace(-pi, pi) at the level size."""
    nx, ny, nz = shape
    vx, vy, vz = voxel_size
    if downsample_res is None:
        downsample_res = shape
    dsx, dsy, dsz = downsample_res
    rx = np.fft.fftshift((np.linspace(-nx // 2, nx // 2, dsx, endpoint=False) * vx).reshape(-1, 1, 1))
    ry = np.fft.fftshift((np.linspace(-ny // 2, ny // 2, dsy, endpoint=False) * vy).reshape(1, -1, 1))
    rz = np.fft.fftshift((np.linspace(-nz // 2, nz // 2, dsz, endpoint=False) * vz).reshape(1, 1, -1))
    kx = np.fft.fftshift(np.linspace(-np.pi, np.pi, dsx, endpoint=False).reshape((-1, 1, 1)))
    ky = np.fft.fftshift(np.linspace(-np.pi, np.pi, dsy, endpoint=False).reshape((1, -1, 1)))
    kz = np.fft.fftshift(np.linspace(-np.pi, np.pi, dsz, endpoint=False).reshape((1, 1, -1)))
    kgrid = {"x": kx, "y": ky, "z": kz}
    rkgrid = {
        "tan": {"x": rz * ky, "y": rx * kz, "z": ry * kx},
        "sin": {"x": ry * kz, "y": rz * kx, "z": rx * ky},
    }
    return kgrid, rkgrid


# ---------------------------------------------------------------------------
# a2-a4: rigid transform and adjoint  (transform.py:5-14, 243-377)
# ---------------------------------------------------------------------------
def shear_factor(rk, theta, mode, inverse=False):
    """transform.py:5-14."""
    if mode == "tan":
        f = np.exp(1j * np.tan(theta / 2) * rk)
    elif mode == "sin":
        f = np.exp(-1j * np.sin(theta) * rk)
    else:
        raise ValueError("Mode must be either 'tan' or 'sin'!")
    return f.conj() if inverse else f


def _steps(params, rkgrid):
    q1, q2, q3, t1, t2, t3 = [float(v) for v in params]
    return (q1, q2, q3), [
        (t1, rkgrid["tan"]["x"], rkgrid["sin"]["x"], 1, 2),
        (t2, rkgrid["tan"]["y"], rkgrid["sin"]["y"], 2, 0),
        (t3, rkgrid["tan"]["z"], rkgrid["sin"]["z"], 0, 1),
    ]


def _shear(arr, axis, phase, w):
    return _ifft(_fft(arr, axis, w) * phase, axis, w)


def translation_factor(kgrid, q1, q2, q3, inverse=False):
    """transform.py:257-261 (complex128, like the reference)."""
    f = np.exp(-1j * (q1 * kgrid["x"] + q2 * kgrid["y"] + q3 * kgrid["z"]))
    return f.conj() if inverse else f


def rigid_apply(x, params, kgrid, rkgrid, inverse=False, prec="f64"):
    """RigidTransformCudaOptimzied.apply, transform.py:263-318.

    T x = Tr(q) Rz(t3) Ry(t2) Rx(t1) x with each R three 1-D Fourier shears
    (tan, sin, tan) and Tr a 3-D Fourier shift.  ``inverse`` only conjugates
    the translation (transform.py:308, quirk)."""
    cdt = _cdt(prec)
    w = WORKERS
    arr = np.array(x, dtype=cdt, copy=True)
    (q1, q2, q3), steps = _steps(params, rkgrid)
    for theta, rkt, rks, at, as_ in steps:
        if prec == "f64":
            vtan = np.exp(1j * np.tan(theta / 2) * rkt)
            vsin = np.exp(-1j * np.sin(theta) * rks)
        else:  # transform.py:275-276: exp evaluated in complex64
            vtan = np.exp((1j * np.tan(theta / 2) * rkt).astype(np.complex64))
            vsin = np.exp((-1j * np.sin(theta) * rks).astype(np.complex64))
        arr = _shear(arr, at, vtan, w).astype(cdt, copy=False)
        arr = _shear(arr, as_, vsin, w).astype(cdt, copy=False)
        arr = _shear(arr, at, vtan, w).astype(cdt, copy=False)
    U = translation_factor(kgrid, q1, q2, q3, inverse)
    arr = _ifftn(_fftn(arr, (0, 1, 2), w) * U, (0, 1, 2), w).astype(cdt, copy=False)
    return arr


def rigid_adjoint(x, params, kgrid, rkgrid, prec="f64"):
    """RigidTransformCudaOptimzied.adjoint, transform.py:320-377: conjugate
    translation first, then rotations z, y, x with conjugated shears."""
    cdt = _cdt(prec)
    w = WORKERS
    arr = np.array(x, dtype=cdt, copy=True)
    (q1, q2, q3), steps = _steps(params, rkgrid)
    U = translation_factor(kgrid, q1, q2, q3, True)
    arr = _ifftn(_fftn(arr, (0, 1, 2), w) * U, (0, 1, 2), w).astype(cdt, copy=False)
    for theta, rkt, rks, at, as_ in steps[::-1]:
        vtan = shear_factor(rkt, theta, "tan", True)
        vsin = shear_factor(rks, theta, "sin", True)
        arr = _shear(arr, at, vtan, w).astype(cdt, copy=False)
        arr = _shear(arr, as_, vsin, w).astype(cdt, copy=False)
        arr = _shear(arr, at, vtan, w).astype(cdt, copy=False)
    return arr


# ---------------------------------------------------------------------------
# a5: derivative operator  (transform.py:379-504)
# ---------------------------------------------------------------------------
def rigid_deriv(x, params, kgrid, rkgrid, p_idx, prec="f64"):
    """RigidTransformDerivativeCuda.apply, transform.py:476-504.

    d/dp of T_d = Tr Rx Ry Rz (rotations applied z, y, x: the REVERSE of
    ``rigid_apply``; quirk Q1).  p<3: translation derivative
    F^H(-i k_p U)F after all rotations (transform.py:422-431); p=3,4,5: the
    three-term product rule over the shears of that rotation
    (transform.py:433-474).  p outside [0,5] returns a copy of the input
    (no else branch, transform.py:476-504)."""
    cdt = _cdt(prec)
    w = WORKERS
    q1, q2, q3, t1, t2, t3 = [float(v) for v in params]
    out = np.array(x, dtype=cdt, copy=True)

    def phases(theta, rkt, rks):
        if prec == "f64":
            vtan = np.exp(1j * np.tan(theta / 2) * rkt)
            vsin = np.exp(-1j * np.sin(theta) * rks)
        else:
            vtan = np.exp((1j * np.tan(theta / 2) * rkt).astype(np.complex64))
            vsin = np.exp((-1j * np.sin(theta) * rks).astype(np.complex64))
        return vtan, vsin

    def rot(a, theta, rkt, rks, at, as_):  # transform.py:392-409
        vtan, vsin = phases(theta, rkt, rks)
        a = _shear(a, at, vtan, w).astype(cdt, copy=False)
        a = _shear(a, as_, vsin, w).astype(cdt, copy=False)
        return _shear(a, at, vtan, w).astype(cdt, copy=False)

    def rot_d(a, theta, rkt, rks, at, as_):  # transform.py:433-474
        vtan, vsin = phases(theta, rkt, rks)
        vtan_d = (1j * ((1 + np.tan(theta / 2) ** 2) / 2) * rkt * vtan).astype(cdt)
        vsin_d = (-1j * np.cos(theta) * rks * vsin).astype(cdt)
        a = _fft(a, at, w).astype(cdt, copy=False)
        s2 = a.copy()
        s3 = a.copy()
        o = _ifft(a * vtan_d, at, w)
        o = _ifft(_fft(o, as_, w) * vsin, as_, w)
        o = _fft(o, at, w) * vtan
        s2 = _ifft(s2 * vtan, at, w)
        s2 = _ifft(_fft(s2, as_, w) * vsin_d, as_, w)
        s2 = _fft(s2, at, w) * vtan
        s3 = _ifft(s3 * vtan, at, w)
        s3 = _ifft(_fft(s3, as_, w) * vsin, as_, w)
        s3 = _fft(s3, at, w) * vtan_d
        return _ifft((o + s2 + s3).astype(cdt, copy=False), at, w).astype(cdt, copy=False)

    def trans(a, deriv_axis=None):  # transform.py:411-431
        U = translation_factor(kgrid, q1, q2, q3)
        if deriv_axis is not None:
            U = -1j * kgrid["xyz"[deriv_axis]] * U
        return _ifftn(_fftn(a, (0, 1, 2), w) * U, (0, 1, 2), w).astype(cdt, copy=False)

    Z = (t3, rkgrid["tan"]["z"], rkgrid["sin"]["z"], 0, 1)
    Y = (t2, rkgrid["tan"]["y"], rkgrid["sin"]["y"], 2, 0)
    X = (t1, rkgrid["tan"]["x"], rkgrid["sin"]["x"], 1, 2)
    if p_idx < 3:
        out = rot(out, *Z)
        out = rot(out, *Y)
        out = rot(out, *X)
        out = trans(out, p_idx)
    elif p_idx == 3:
        out = rot(out, *Z)
        out = rot(out, *Y)
        out = rot_d(out, *X)
        out = trans(out)
    elif p_idx == 4:
        out = rot(out, *Z)
        out = rot_d(out, *Y)
        out = rot(out, *X)
        out = trans(out)
    elif p_idx == 5:
        out = rot_d(out, *Z)
        out = rot(out, *Y)
        out = rot(out, *X)
        out = trans(out)
    return out


def rigid_deriv_apply(x, params, kgrid, rkgrid, p_idx):
    """d/dtheta_p of ``rigid_apply``'s T = Tr Rz Ry Rx (rotations applied x,
    y, z: transform.py:267-318), in FP64: the exact Jacobian column of
    RigidTransformCudaOptimzied.apply, as needed by the gradient of a loss
    through ``apply``.  Same building blocks as the reference's derivative
    class (translation derivative -i k_p U, transform.py:422-431; three-term
    product rule over the shears, transform.py:433-474), in the apply order."""
    w = WORKERS
    q1, q2, q3, t1, t2, t3 = [float(v) for v in params]
    a = np.array(x, dtype=np.complex128, copy=True)

    def rot(a, theta, rkt, rks, at, as_, deriv):
        vtan = np.exp(1j * np.tan(theta / 2) * rkt)
        vsin = np.exp(-1j * np.sin(theta) * rks)
        if not deriv:
            return _shear(_shear(_shear(a, at, vtan, w), as_, vsin, w), at, vtan, w)
        vtan_d = 1j * ((1 + np.tan(theta / 2) ** 2) / 2) * rkt * vtan
        vsin_d = -1j * np.cos(theta) * rks * vsin
        return (_shear(_shear(_shear(a, at, vtan_d, w), as_, vsin, w), at, vtan, w)
                + _shear(_shear(_shear(a, at, vtan, w), as_, vsin_d, w), at, vtan, w)
                + _shear(_shear(_shear(a, at, vtan, w), as_, vsin, w), at, vtan_d, w))

    a = rot(a, t1, rkgrid["tan"]["x"], rkgrid["sin"]["x"], 1, 2, p_idx == 3)
    a = rot(a, t2, rkgrid["tan"]["y"], rkgrid["sin"]["y"], 2, 0, p_idx == 4)
    a = rot(a, t3, rkgrid["tan"]["z"], rkgrid["sin"]["z"], 0, 1, p_idx == 5)
    U = translation_factor(kgrid, q1, q2, q3)
    if p_idx < 3:
        U = -1j * kgrid["xyz"[p_idx]] * U
    return _ifftn(_fftn(a, (0, 1, 2), w) * U, (0, 1, 2), w)


def dc_loss(x, transforms, ksp, mps, shot_mask, kgrid, rkgrid):
    """Data-consistency loss 0.5 sum_s ||M_s (F(S T_s x) - y)||^2 in FP64 (the
    per-shot residual of transform_estimation.py:41-43, uncentred frame, T
    the ``apply`` of transform.py:263-318)."""
    tot = 0.0
    for s in range(transforms.shape[0]):
        r = shot_residual(ksp, mps, shot_mask[s], x, transforms[s], kgrid, rkgrid)
        tot += 0.5 * float(np.vdot(r, r).real)
    return tot


def dc_loss_grad(x, transforms, ksp, mps, shot_mask, kgrid, rkgrid):
    """Gradients of ``dc_loss`` in FP64: with r_s the masked residual,
    d/dRe(x) + i d/dIm(x) = sum_s T_s^H sum_c conj(S_c) F^H r_s, and
    d/dtheta_{s,p} = Re <r_s, M_s F(S dT_s/dtheta_p x)>."""
    gx = np.zeros(np.shape(x), dtype=np.complex128)
    gt = np.zeros(np.shape(transforms), dtype=np.float64)
    for s in range(transforms.shape[0]):
        r = shot_residual(ksp, mps, shot_mask[s], x, transforms[s], kgrid, rkgrid)
        gx += encode_adjoint(r, mps, shot_mask[s], transforms[s], kgrid, rkgrid)
        for p in range(6):
            col = rigid_deriv_apply(x, transforms[s], kgrid, rkgrid, p)
            jc = _fftn(mps * col, (-3, -2, -1), WORKERS) * shot_mask[s]
            gt[s, p] = float(np.vdot(r, jc).real)
    return gx, gt


def rigid_apply_reverse(x, params, kgrid, rkgrid, prec="f64"):
    """T_d x = Tr Rx Ry Rz x: the composition whose Jacobian ``rigid_deriv``
    returns (rotations z, y, x as in transform.py:481-483, then
    transform.py:411-420).  Used by the finite-difference pin."""
    cdt = _cdt(prec)
    w = WORKERS
    q1, q2, q3, t1, t2, t3 = [float(v) for v in params]
    a = np.array(x, dtype=cdt, copy=True)
    for theta, rkt, rks, at, as_ in [
        (t3, rkgrid["tan"]["z"], rkgrid["sin"]["z"], 0, 1),
        (t2, rkgrid["tan"]["y"], rkgrid["sin"]["y"], 2, 0),
        (t1, rkgrid["tan"]["x"], rkgrid["sin"]["x"], 1, 2),
    ]:
        vtan = np.exp(1j * np.tan(theta / 2) * rkt)
        vsin = np.exp(-1j * np.sin(theta) * rks)
        a = _shear(a, at, vtan, w)
        a = _shear(a, as_, vsin, w)
        a = _shear(a, at, vtan, w)
    U = translation_factor(kgrid, q1, q2, q3)
    return _ifftn(_fftn(a, (0, 1, 2), w) * U, (0, 1, 2), w).astype(cdt, copy=False)


# ---------------------------------------------------------------------------
# encoding helpers (inlined in the reference: image_estimation.py:93-129,
# transform_estimation.py:41-55, joint_recon.py:111-113)
# ---------------------------------------------------------------------------
def _vdot_real(a, b, prec):
    if prec == "f64":
        return float(np.vdot(a.astype(np.complex128, copy=False), b.astype(np.complex128, copy=False)).real)
    return float(np.float32(np.vdot(a, b).real))


def encode_forward(x, mps, mask_s, params, kgrid, rkgrid, prec="f64"):
    """mask_s * fftn(mps * T x)  (transform_estimation.py:42-45 without the
    subtraction; joint_recon.py:112-113)."""
    cdt = _cdt(prec)
    Tx = rigid_apply(x, params, kgrid, rkgrid, prec=prec)
    return (_fftn((mps * Tx).astype(cdt, copy=False), (-3, -2, -1), WORKERS) * mask_s).astype(cdt, copy=False)


def encode_adjoint(y, mps, mask_s, params, kgrid, rkgrid, prec="f64"):
    """T^H sum_c conj(mps) ifftn(mask_s * y)  (image_estimation.py:94-102)."""
    cdt = _cdt(prec)
    t = _ifftn((mask_s * y).astype(cdt, copy=False), (-3, -2, -1), WORKERS)
    return rigid_adjoint(np.sum(np.conj(mps) * t, axis=0), params, kgrid, rkgrid, prec=prec)


def cg_rhs(ksp, mps, shot_mask, transforms, kgrid, rkgrid, M, prec="f64"):
    """image_estimation.py:92-103."""
    cdt = _cdt(prec)
    b = np.zeros(ksp.shape[1:], dtype=cdt)
    for i in range(transforms.shape[0]):
        temp = _ifftn((shot_mask[i] * ksp).astype(cdt, copy=False), (-3, -2, -1), WORKERS)
        b += rigid_adjoint(np.sum(np.conj(mps) * temp, axis=0), transforms[i], kgrid, rkgrid, prec=prec)
        b *= M
    return b


def normal_op(x, mps, shot_mask, transforms, kgrid, rkgrid, M, prec="f64"):
    """A_func, image_estimation.py:106-129: M * sum_s T_s^H sum_c conj(S_c)
    F^H M_s F S_c T_s x."""
    cdt = _cdt(prec)
    out = np.zeros(x.shape, dtype=cdt)
    for i in range(transforms.shape[0]):
        Ex = rigid_apply(x, transforms[i], kgrid, rkgrid, prec=prec)
        Ex = (mps * Ex).astype(cdt, copy=False)
        Ex = _fftn(Ex, (-3, -2, -1), WORKERS) * shot_mask[i]
        Ex = _ifftn(Ex.astype(cdt, copy=False), (-3, -2, -1), WORKERS)
        out += rigid_adjoint(np.sum(np.conj(mps) * Ex, axis=0), transforms[i], kgrid, rkgrid, prec=prec)
    out *= M
    return out


def estimate_image_cg(ksp, mps, shot_mask, transforms, kgrid, rkgrid, P, M, x0=None, max_iter=100,
                      tol=1e-6, prec="f64", history=None):
    """image_estimation.py:10-165 (PCG).  ``history``, if a list, receives
    (alpha, beta, rsnew) per iteration."""
    cdt = _cdt(prec)
    x = np.zeros(ksp.shape[1:], dtype=cdt) if x0 is None else np.array(x0, dtype=cdt, copy=True)
    b = cg_rhs(ksp, mps, shot_mask, transforms, kgrid, rkgrid, M, prec)
    Ap = normal_op(x, mps, shot_mask, transforms, kgrid, rkgrid, M, prec)
    r = b - Ap
    z = (P * r).astype(cdt, copy=False)
    p = z.copy()
    rsold = _vdot_real(r, z, prec)
    for _ in range(max_iter):
        Ap = normal_op(p, mps, shot_mask, transforms, kgrid, rkgrid, M, prec)
        pAp = _vdot_real(p, Ap, prec)
        alpha = rsold / pAp
        x = (x + alpha * p).astype(cdt, copy=False)
        r = (r - alpha * Ap).astype(cdt, copy=False)
        z = (P * r).astype(cdt, copy=False)
        rsnew = _vdot_real(r, z, prec)
        if history is not None:
            history.append((alpha, rsnew / rsold, rsnew))
        if math.sqrt(abs(rsnew)) < tol:
            break
        beta = rsnew / rsold
        p = (z + beta * p).astype(cdt, copy=False)
        rsold = rsnew
    return x


# ---------------------------------------------------------------------------
# a9-a10: LM motion update  (transform_estimation.py:7-152, use_ga=False)
# ---------------------------------------------------------------------------
def shot_residual(ksp, mps, mask_s, image, params, kgrid, rkgrid, prec="f64"):
    """transform_estimation.py:41-45."""
    cdt = _cdt(prec)
    Tx = rigid_apply(image, params, kgrid, rkgrid, prec=prec)
    r = _fftn((mps * Tx).astype(cdt, copy=False), (-3, -2, -1), WORKERS)
    return ((r - ksp) * mask_s).astype(cdt, copy=False)


def shot_jacobian(mps, mask_s, image, params, kgrid, rkgrid, prec="f64"):
    """transform_estimation.py:51-55: six columns mask*fftn(mps*dT_p x)."""
    cdt = _cdt(prec)
    cols = []
    for p_idx in range(6):
        d = rigid_deriv(image, params, kgrid, rkgrid, p_idx, prec=prec)
        cols.append((_fftn((mps * d).astype(cdt, copy=False), (-3, -2, -1), WORKERS) * mask_s).astype(cdt, copy=False))
    return cols


def lm_normal_equations(residual, cols, damp_s, prec="f64"):
    """transform_estimation.py:47-68: e0, gradient, damped Gram matrix."""
    e0 = _vdot_real(residual, residual, prec)
    g = np.array([_vdot_real(residual, c, prec) for c in cols], dtype=np.float64)
    H = np.zeros((6, 6), dtype=np.float64)
    for i in range(6):
        for j in range(i, 6):
            v = _vdot_real(cols[i], cols[j], prec)
            if i == j:
                H[i, j] = v + damp_s
            else:
                H[i, j] = v
                H[j, i] = v
    return e0, g, H


GA_EPS = 1e-4  # transform_estimation.py:76


def ga_direction(mps, mask_s, image, params, cols, v, kgrid, rkgrid, prec="f64"):
    """transform_estimation.py:78-100: J'v ~ (J(theta + eps v) v - J(theta) v) / eps
    with the LM's (reverse-order) Jacobian columns."""
    cdt = _cdt(prec)
    jv = sum(v[p] * cols[p] for p in range(6))
    cols_eps = shot_jacobian(mps, mask_s, image, params + GA_EPS * v, kgrid, rkgrid, prec)
    jev = sum(v[p] * cols_eps[p] for p in range(6))
    return ((jev - jv) / GA_EPS).astype(cdt, copy=False)


def ga_step(H, cols, jpv, v, prec="f64"):
    """transform_estimation.py:102-117: rhs_p = -Re<J'v, J_p>, a = H^-1 rhs,
    v + a/2 if |a| <= |v| else v."""
    rhs = np.array([-_vdot_real(jpv, cols[p], prec) for p in range(6)], dtype=np.float64)
    a = np.linalg.solve(H, rhs)
    return v + 0.5 * a if np.linalg.norm(a) <= 1.0 * np.linalg.norm(v) else v


def estimate_transform(ksp, mps, shot_mask, image, kgrid, rkgrid, damp, convergence_flags, transforms,
                       tol=1e-6, prec="f64", trace=None, use_ga=False):
    """transform_estimation.py:7-152.

    Mutates ``transforms``, ``damp`` and ``convergence_flags`` in place and
    returns T(mean) image.  ``trace``, if a list, receives per-shot
    (e0, e1, accepted, delta).  ``use_ga`` adds the reference's experimental
    geodesic acceleration (transform_estimation.py:73-121)."""
    nshots = shot_mask.shape[0]
    for s in range(nshots):
        res = shot_residual(ksp, mps, shot_mask[s], image, transforms[s], kgrid, rkgrid, prec)
        cols = shot_jacobian(mps, shot_mask[s], image, transforms[s], kgrid, rkgrid, prec)
        e0, g, H = lm_normal_equations(res, cols, damp[s], prec)
        delta = np.linalg.solve(H, g)  # transform_estimation.py:71
        if use_ga:
            v = -delta
            jpv = ga_direction(mps, shot_mask[s], image, transforms[s], cols, v, kgrid, rkgrid, prec)
            nxt = transforms[s] + ga_step(H, cols, jpv, v, prec)
        else:
            nxt = transforms[s] - delta  # transform_estimation.py:125
        res1 = shot_residual(ksp, mps, shot_mask[s], image, nxt, kgrid, rkgrid, prec)
        e1 = _vdot_real(res1, res1, prec)
        acc = e1 < e0
        if acc:  # transform_estimation.py:138-144
            if np.all(np.abs(transforms[s] - nxt) < tol):
                convergence_flags[s] = True
            transforms[s] = nxt.copy()
            damp[s] = max(damp[s] / 5, 1e-4)
        else:
            damp[s] = min(damp[s] * 1.5, 1e16)
        if trace is not None:
            trace.append((e0, e1, bool(acc), delta.copy()))
    mean_t = np.mean(transforms, axis=0)  # transform_estimation.py:147-150
    out = rigid_apply(image, mean_t, kgrid, rkgrid, prec=prec)
    transforms -= mean_t
    return out


# ---------------------------------------------------------------------------
# uncorrected SENSE baseline: sigpy.mri.app.SenseRecon (simulation.py:462-466,
# recon.py:153).  SigPy is not installed; its published algorithm is
# restated: LinearLeastSquares with ConjugateGradient on
# (A^H A + lamda I) x = A^H W^1/2 y, A = W^1/2 F_c S, F_c the centred ortho FFT.
# ---------------------------------------------------------------------------
def _fftc(a, axes):
    return np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(a, axes=axes), axes=axes, norm="ortho"), axes=axes)


def _ifftc(a, axes):
    return np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(a, axes=axes), axes=axes, norm="ortho"), axes=axes)


def sense_recon(y, mps, lamda=0.0, weights=None, P=None, max_iter=100, tol=0.0, x0=None, info=None):
    """sigpy.mri.app.SenseRecon(y, mps, lamda, weights, P=..., max_iter, tol).run()
    in FP64.  weights None -> rss(y over coils) > 0 (sigpy.mri.app
    _estimate_weights).  The CG is sigpy.alg.ConjugateGradient: stop when
    sqrt(<r, P r>) <= tol (also before the first step), or on <p, A p> <= 0
    without updating; the last iteration updates x only."""
    ax = (-3, -2, -1)
    if weights is None:
        weights = (np.sqrt(np.sum(np.abs(y) ** 2, axis=0)) > 0).astype(np.float64)
    sw = np.sqrt(weights)

    def AHA(v):
        k = sw * _fftc(mps * v, ax)
        return np.sum(np.conj(mps) * _ifftc(sw * k, ax), axis=0) + lamda * v

    b = np.sum(np.conj(mps) * _ifftc(sw * (sw * y), ax), axis=0)
    x = np.zeros(y.shape[1:], dtype=np.complex128) if x0 is None else np.array(x0, dtype=np.complex128)
    Pm = (lambda v: v) if P is None else (lambda v: P * v)
    r = b - AHA(x)
    z = Pm(r)
    p = z.copy()
    rzold = float(np.vdot(r, z).real)
    resid = math.sqrt(abs(rzold))
    it = 0
    while not (it >= max_iter or resid <= tol):
        Ap = AHA(p)
        pAp = float(np.vdot(p, Ap).real)
        if pAp <= 0:
            break
        alpha = rzold / pAp
        x = x + alpha * p
        if it < max_iter - 1:
            r = r - alpha * Ap
            z = Pm(r)
            rznew = float(np.vdot(r, z).real)
            p = z + (rznew / rzold) * p
            rzold = rznew
        resid = math.sqrt(abs(rzold))
        it += 1
    if info is not None:
        info["iterations"] = it
    return x


# ---------------------------------------------------------------------------
# calibration preprocessing (noise_whiten.py:5-14, coil_compression.py:6-46)
# ---------------------------------------------------------------------------
def noise_whiten(ksp, noise, reg=1e-6):
    """noise_whiten.py:5-14 (FP64): W = inv(chol(cov + reg I))^H applied to
    the coil axis."""
    C = ksp.shape[0]
    nz = np.asarray(noise, dtype=np.complex128).reshape(C, -1)
    cov = (nz @ nz.conj().T) / nz.shape[1] + reg * np.eye(C)
    L = np.linalg.cholesky(cov)
    W = np.linalg.inv(L).conj().T
    return (W @ np.asarray(ksp, dtype=np.complex128).reshape(C, -1)).reshape(ksp.shape), W


def coil_compression(mps, ksp=None, perc=0.95, reg=1e-6):
    """coil_compression.py:6-46 (FP64): P = (S norm) S^H, U S V^H = svd(P),
    r = argmax(cumsum(S)/sum(S) >= perc), A = U^H[:r].  Returns
    (cc_mps, cc_ksp, A, S)."""
    m = np.asarray(mps, dtype=np.complex128)
    C = m.shape[0]
    norm = 1 / (np.sum(m * m.conj(), axis=0).real + reg)
    mf = m.reshape(C, -1)
    P = (mf * norm.ravel()) @ mf.conj().T
    U, S, _ = np.linalg.svd(P)
    energy = np.cumsum(S) / np.sum(S)
    r = int(np.argmax(energy >= perc))
    A = U.conj().T[:r]
    cc_mps = (A @ mf).reshape(r, *m.shape[1:])
    cc_ksp = None if ksp is None else (A @ np.asarray(ksp, np.complex128).reshape(C, -1)).reshape(r, *m.shape[1:])
    return cc_mps, cc_ksp, A, S


def hankel_fwd(x, kernel_size):
    """espirit.py:9-40: (kx*ky*kz, nblocks, N1) with blocks ordered kx_start
    fastest, each block flattened in Fortran order."""
    Nx, Ny, Nz, N1 = x.shape
    kx, ky, kz = kernel_size
    nbx, nby, nbz = Nx - kx + 1, Ny - ky + 1, Nz - kz + 1
    h = np.zeros((kx * ky * kz, nbx * nby * nbz, N1), dtype=np.complex128)
    idx = 0
    for z0 in range(nbz):
        for y0 in range(nby):
            for x0 in range(nbx):
                h[:, idx, :] = x[x0:x0 + kx, y0:y0 + ky, z0:z0 + kz, :].reshape((kx * ky * kz, N1), order="F")
                idx += 1
    return h


def fold_rx(h):
    """espirit.py:42-45: (K, nb, N1) -> (K*N1, nb), row = k + K*c."""
    hp = np.transpose(h, (0, 2, 1))
    return hp.reshape((-1, hp.shape[2]), order="F")


def rx_espirit_3d(calib, imsize, kernel_size=(5, 5, 5), eig_thresh=0.02, mask_thresh=0.99, info=None, z_slices=None):
    """espirit.py:47-239 in FP64: calib (NKx, NKy, NKz, NRx) -> maps (Nx, Ny, Nz, NRx).
    ``z_slices`` (timing samples only) restricts the per-voxel stage to those
    z indices; the other slices stay zero."""
    calib = np.asarray(calib, dtype=np.complex128)
    NRx = calib.shape[3]
    kx, ky, kz = kernel_size
    sens = np.zeros((*imsize, NRx), dtype=np.complex128)
    H = fold_rx(hankel_fwd(calib, kernel_size))
    U, S, _ = np.linalg.svd(H, full_matrices=False)
    if eig_thresh < 1:
        U = U[:, S > S[0] * eig_thresh] if S.size and S[0] > 0 else U[:, :0]
    else:
        U = U[:, :int(round(eig_thresh))]
    if info is not None:
        info["kept"] = U.shape[1]
        info["singular_values"] = S
    if U.shape[1] == 0:
        return sens
    nc = U.shape[1]
    Ur = U.reshape((kx, ky, kz, NRx, nc), order="F")
    pz0, pz1 = int(math.ceil((imsize[2] - kz) / 2)), int(math.floor((imsize[2] - kz) / 2))
    Uz = np.pad(Ur, [(0, 0), (0, 0), (pz0, pz1), (0, 0), (0, 0)])
    Uz = np.fft.fftshift(np.fft.ifft(np.fft.ifftshift(Uz, axes=2), axis=2), axes=2)
    px0, px1 = int(math.ceil((imsize[0] - kx) / 2)), int(math.floor((imsize[0] - kx) / 2))
    py0, py1 = int(math.ceil((imsize[1] - ky) / 2)), int(math.floor((imsize[1] - ky) / 2))
    norm_factor = np.prod(imsize) / np.sqrt(np.prod(kernel_size))
    lam_max = np.zeros(imsize)
    for z in (range(imsize[2]) if z_slices is None else z_slices):
        t = np.pad(Uz[:, :, z], [(px0, px1), (py0, py1), (0, 0), (0, 0)])
        t = np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(t, axes=(0, 1)), axes=(0, 1)), axes=(0, 1))
        M = t.reshape(-1, NRx, nc)
        C = M @ np.conj(np.transpose(M, (0, 2, 1)))
        w, v = np.linalg.eigh(C)
        imax = np.argmax(w.real, axis=1)
        pv = np.take_along_axis(v, imax[:, None, None], axis=2)[:, :, 0]
        mask = np.sqrt(np.maximum(w[:, -1], 0)) * norm_factor > mask_thresh
        lam_max[:, :, z] = (np.sqrt(np.maximum(w[:, -1], 0)) * norm_factor).reshape(imsize[0], imsize[1])
        sl = np.zeros((imsize[0] * imsize[1], NRx), dtype=np.complex128)
        sl[mask] = pv[mask]
        sens[:, :, z] = sl.reshape(imsize[0], imsize[1], NRx)
    if info is not None:
        info["mask_value"] = lam_max  # sqrt(lambda_max) * norm_factor per voxel
    return sens * np.exp(-1j * np.angle(sens[..., 0:1]))


# ---------------------------------------------------------------------------
# a11-a12: JointRecon  (joint_recon.py:38-115)
# ---------------------------------------------------------------------------
class JointRecon:
    """joint_recon.py:10-115 on the CPU.  Inputs are centred (the class
    uncentres them, joint_recon.py:63-65)."""

    def __init__(self, ksp, mps, shot_mask, kgrid, rkgrid, t0=None, max_cg_iter=3, max_nm_iter=1,
                 max_joint_iter=100, prec="f64", compute_objective=True):
        cdt = _cdt(prec)
        self.prec = prec
        self.kgrid, self.rkgrid = kgrid, rkgrid
        self.max_cg_iter, self.max_nm_iter, self.max_joint_iter = max_cg_iter, max_nm_iter, max_joint_iter
        self.compute_objective = compute_objective
        self.num_shots = len(shot_mask)
        self.img_shape = ksp.shape[1:]
        self.ksp = np.fft.ifftshift(np.asarray(ksp), axes=(-3, -2, -1)).astype(cdt)
        self.mps = np.fft.ifftshift(np.asarray(mps), axes=(-3, -2, -1)).astype(cdt)
        self.shot_mask = np.fft.ifftshift(np.asarray(shot_mask), axes=(-3, -2, -1))
        self.max_norm = float(np.max(np.abs(_ifftn(self.ksp, (-3, -2, -1), WORKERS))))
        self.ksp /= self.max_norm
        self.x = np.zeros(self.img_shape, dtype=cdt)
        self.t = np.zeros((self.num_shots, 6), dtype=np.float64) if t0 is None else np.array(t0, dtype=np.float64)
        self.damp = np.ones(self.num_shots)
        self.convergence_flags = np.zeros(self.num_shots, dtype=bool)
        ssq = np.sum(np.abs(self.mps) ** 2, axis=0)
        self.P = (1 / (ssq + 1e-3)).astype(_rdt(prec))
        self.M = ssq > 0.1

    def step(self):
        """One outer iteration of joint_recon.py:92-102."""
        self.x = estimate_image_cg(self.ksp, self.mps, self.shot_mask, self.t, self.kgrid, self.rkgrid,
                                   self.P, self.M, x0=self.x, max_iter=self.max_cg_iter, prec=self.prec)
        for _ in range(self.max_nm_iter):
            self.x = estimate_transform(self.ksp, self.mps, self.shot_mask, self.x, self.kgrid, self.rkgrid,
                                        self.damp, self.convergence_flags, self.t, prec=self.prec)
            self.x *= self.M

    def objective(self):
        """joint_recon.py:108-115."""
        total = 0.0
        for s in range(self.num_shots):
            r = shot_residual(self.ksp, self.mps, self.shot_mask[s], self.x, self.t[s], self.kgrid,
                              self.rkgrid, self.prec)
            total += _vdot_real(r, r, self.prec)
        return total

    def run(self):
        self.transform_history, self.objective_history = [], []
        for _ in range(self.max_joint_iter):
            self.step()
            self.transform_history.append(self.t.copy())
            if self.compute_objective:
                self.objective_history.append(self.objective())
        self.transform_history = np.stack(self.transform_history)
        self.objective_history = np.array(self.objective_history)
        return np.fft.fftshift(self.x * self.max_norm), self.t


# ---------------------------------------------------------------------------
# a13: pyramid  (pyramid.py:24-118) with SigPy's centred fft / resize
# ---------------------------------------------------------------------------
def sp_resize(a, oshape):
    """SigPy ``resize`` semantics: centred crop/zero-pad per axis
    (ishift = max(i//2 - o//2, 0), oshift = max(o//2 - i//2, 0))."""
    a = np.asarray(a)
    out = np.zeros(oshape, dtype=a.dtype)
    src, dst = [], []
    for i, o in zip(a.shape, oshape):
        ishift = max(i // 2 - o // 2, 0)
        oshift = max(o // 2 - i // 2, 0)
        n = min(i - ishift, o - oshift)
        src.append(slice(ishift, ishift + n))
        dst.append(slice(oshift, oshift + n))
    out[tuple(dst)] = a[tuple(src)]
    return out


def sp_fft(a, axes):
    """SigPy centred orthonormal FFT over ``axes``."""
    return np.fft.fftshift(_fftn(np.fft.ifftshift(a, axes=axes), axes, WORKERS), axes=axes)


def sp_ifft(a, axes):
    """SigPy centred orthonormal inverse FFT over ``axes``."""
    return np.fft.fftshift(_ifftn(np.fft.ifftshift(a, axes=axes), axes, WORKERS), axes=axes)


def pyramid_levels(ksp, mps, A, nshots, n_spatial_levels=3, sampling="disorder"):
    """pyramid.py:55-72: per coarse level, (down_factor, ksp_res, mps_res,
    A_res, ds_shape)."""
    full = ksp.shape[1:]
    ncoils = ksp.shape[0]
    out = []
    for sl in range(n_spatial_levels - 1):
        f = 2 ** (n_spatial_levels - sl - 1)
        if sampling == "disorder":
            ds = (int(full[0] // f), int(full[1] // f), full[2])
        else:
            ds = (int(full[0] // f), full[1], full[2])
        mps_res = sp_ifft(sp_resize(sp_fft(mps, (-3, -2, -1)), (ncoils, *ds)), (-3, -2, -1))
        ksp_res = sp_resize(ksp, (ncoils, *ds))
        A_res = sp_resize(A, (nshots, *ds[:2], 1))
        out.append((f, ksp_res, mps_res, A_res, ds))
    return out


def pyramid_reconstruction(ksp, mps, A, nshots, n_spatial_levels=3, n_joint_iters=100, sampling="disorder",
                           final_cg_iter=100):
    """pyramid.py:24-118: JointRecon on each coarse level (t carried across
    levels unscaled, quirk Q13), then CG only at full resolution with the
    estimated motion (pyramid.py:96-115).  Returns (fftshift(x*max_norm), t)."""
    full = ksp.shape[1:]
    t_est = np.zeros((nshots, 6), dtype=np.float64)
    for f, ksp_res, mps_res, A_res, ds in pyramid_levels(ksp, mps, A, nshots, n_spatial_levels, sampling):
        kg, rkg = compute_transform_grids_voxel(full, (1, 1, 1), ds)
        app = JointRecon(ksp_res, mps_res, A_res, kg, rkg, t0=t_est, max_joint_iter=int(n_joint_iters * f),
                         compute_objective=False)
        _, t_est = app.run()
        t_est = np.array(t_est)
    k = np.fft.ifftshift(np.asarray(ksp), axes=(-3, -2, -1)).astype(np.complex128)
    m = np.fft.ifftshift(np.asarray(mps), axes=(-3, -2, -1)).astype(np.complex128)
    a = np.fft.ifftshift(np.asarray(A), axes=(-3, -2, -1)).astype(bool)
    max_norm = float(np.max(np.abs(_ifftn(k, (-3, -2, -1), WORKERS))))
    k /= max_norm
    kg, rkg = compute_transform_grids_voxel(full, (1, 1, 1))
    ssq = np.sum(np.abs(m) ** 2, axis=0)
    x = estimate_image_cg(k, m, a, t_est, kg, rkg, 1 / (ssq + 1e-3), ssq > 0.1, x0=np.zeros(full, np.complex128),
                          max_iter=final_cg_iter)
    return np.fft.fftshift(x * max_norm), t_est


# ---------------------------------------------------------------------------
# a14-a15: DISORDER sampling and the corrupted forward model
# (simulation.py:65-85, 192-245, 383-410)
# ---------------------------------------------------------------------------
@dataclass
class DisorderConfig:
    """simulation.py:192-199."""
    max_line_number: int
    max_partition_number: int
    pat_lines_to_measure: int
    block_lin: int
    block_par: int
    seed: int = 123456


def make_disorder_order(cfg: DisorderConfig) -> Tuple[List[Tuple[int, int]], Tuple[int, int]]:
    """simulation.py:201-245: tiles of (block_lin x block_par), one stdlib
    ``random.Random(seed)`` shuffle per tile in (linG, parG) order, the
    first ``indices_per_group`` draws interleaved across tiles."""
    L = math.ceil((cfg.max_line_number + 1) / cfg.block_lin) * cfg.block_lin
    P = math.ceil((cfg.max_partition_number + 1) / cfg.block_par) * cfg.block_par
    max_line_number = L - 1
    max_partition_number = P - 1
    n_reorder = cfg.pat_lines_to_measure * (max_partition_number + 1)
    per_group_max = cfg.block_lin * cfg.block_par
    n_lin_g = L // cfg.block_lin
    n_par_g = P // cfg.block_par
    n_groups = n_lin_g * n_par_g
    per_group = n_reorder // n_groups
    rng = random.Random(cfg.seed)
    base = list(range(per_group_max))
    out = [None] * (per_group * n_groups)
    for linG in range(n_lin_g):
        for parG in range(n_par_g):
            idx = base[:]
            rng.shuffle(idx)
            for j in range(per_group):
                local = idx[j]
                ll = local % cfg.block_lin
                lp = local // cfg.block_lin
                line = ll + (linG % cfg.block_par) * (max_line_number + 1) // cfg.block_par + (linG // cfg.block_par) * cfg.block_lin
                part = lp + parG * cfg.block_par
                out[j * n_groups + linG * n_par_g + parG] = (line, part)
    return out, (L, P)


def disorder_masks(order, nshots, K):
    """simulation.py:396-402: contiguous chunks of ``len(order)//nshots``
    samples per shot; mask[s, partition, line] = True; trailing samples
    dropped.  Returns (nshots, K0, K1, 1) bool (simulation.py:437)."""
    order = np.asarray(order, dtype=int)
    A = np.zeros((nshots, *K), dtype=bool)
    sps = order.shape[0] // nshots
    for s in range(nshots):
        A[s, order[s * sps:(s + 1) * sps, 1], order[s * sps:(s + 1) * sps, 0]] = True
    return A[..., None]


def generate_corrupted_kspace(gt_img, mps, A, transforms, prec="f64"):
    """simulation.py:65-85: centred in, centred out."""
    cdt = _cdt(prec)
    gt = np.fft.ifftshift(gt_img)
    m = np.fft.ifftshift(mps, axes=(-3, -2, -1))
    a = np.fft.ifftshift(A, axes=(-3, -2, -1))
    kgrid, rkgrid = compute_transform_grids_voxel(gt.shape, [1, 1, 1])
    ksp = np.zeros(m.shape, dtype=cdt)
    for s in range(transforms.shape[0]):
        t = rigid_apply(gt, transforms[s], kgrid, rkgrid, prec=prec)
        ksp += (_fftn((t * m).astype(cdt, copy=False), (-3, -2, -1), WORKERS) * a[s]).astype(cdt, copy=False)
    return np.fft.fftshift(ksp, axes=(-3, -2, -1))


# ---------------------------------------------------------------------------
# f4: metrics and shot masks of the real-data scripts
# (simulation.py:87-97, 476-477; recon.py:158-165)
# ---------------------------------------------------------------------------
def gradient_entropy(vol):
    """simulation.py:87-97: forward differences per axis with the last one
    zero, gradient magnitude in the volume's dtype, NumPy's 256-bin density
    histogram over [0, max + 1e-8], entropy of (h + 1e-12) in bits.  Returns
    (entropy, counts)."""
    v = np.asarray(vol)
    g = [np.diff(v, axis=a, append=np.take(v, [-1], axis=a)) for a in range(3)]
    gmag = np.sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2])
    counts, edges = np.histogram(gmag, bins=256, range=(0, gmag.max() + 1e-8))
    dens = counts / np.diff(edges) / counts.sum()
    h = dens + 1e-12
    return float(-np.sum(h * np.log2(h))), counts


def error_volume(corrected, gt):
    """simulation.py:476-477: |corrected - gt| and its 2-norm."""
    e = np.abs(np.asarray(corrected) - np.asarray(gt))
    return e, float(np.linalg.norm(e))


def shot_masks_from_counters(lines, partitions, n_shots, n_par, n_lin):
    """recon.py:158-165, loop form: shot s marks (partition, line) of samples
    s*sps .. (s+1)*sps - 1, sps = len(lines) // n_shots."""
    mask = np.zeros((n_shots, n_par, n_lin), dtype=bool)
    sps = len(lines) // n_shots
    for s in range(n_shots):
        for j in range(s * sps, (s + 1) * sps):
            mask[s, partitions[j], lines[j]] = True
    return mask[..., None]
