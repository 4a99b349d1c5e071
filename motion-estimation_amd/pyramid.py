"""Multi-resolution driver (pyramid.py:24-118) on the GPU.

Coarse levels crop k-space (and the maps' spectra) to (N0/f, N1/f, N2) with
SigPy's centred FFT / centred resize semantics (pyramid.py:65-67) and run a
JointRecon with f * n_joint_iters outer iterations, carrying the motion
estimates across levels without rescaling (pyramid.py:73-74).  The final
level is image-only CG at full resolution with the reference's default 100
iterations (pyramid.py:96-115)."""
from __future__ import annotations

import os

import numpy as np
import torch

from . import _lib
from .encoding import SenseProblem, permute, shift_offsets, std_strides
from .joint_recon import JointRecon
from .utils import as_c64, compute_transform_grids_voxel


def _fft3(x, shape, ncoils, inverse):
    n0, n1, n2 = shape
    vol = n0 * n1 * n2
    for ax in range(3):
        _lib.call("me_fft_axis", _lib.ptr(x), vol, _lib.ptr(x), vol, ncoils, _lib.int3(shape),
                  _lib.ll3((n1 * n2, n2, 1)), ax, int(inverse), _lib.stream(x.device))
    return x


def _shift(x, shape, ncoils, kind):
    out = torch.empty_like(x)
    dims = (ncoils, *shape)
    permute(x, std_strides(0, shape), dims, out, std_strides(0, shape), dims, shift_offsets(shape, kind))
    return out


def centered_crop(x, oshape):
    """SigPy ``resize`` to a smaller centred window, per axis
    ishift = n//2 - m//2 (sizes of (ncoils, N0, N1, N2) arrays)."""
    ncoils = x.shape[0]
    ishape = tuple(x.shape[1:])
    out = torch.empty((ncoils, *oshape), dtype=torch.complex64, device=x.device)
    off = (0, *[n // 2 - m // 2 for n, m in zip(ishape, oshape)])
    permute(x, std_strides(0, ishape), (ncoils, *ishape), out, std_strides(0, oshape), (ncoils, *oshape), off)
    return out


def resample_maps(mps, oshape):
    """sp.ifft(sp.resize(sp.fft(mps, axes), (C, *oshape)), axes) (pyramid.py:65):
    centred orthonormal FFT, centred crop of the spectrum, centred inverse FFT."""
    ncoils = mps.shape[0]
    ishape = tuple(mps.shape[1:])
    spec = _fft3(_shift(mps, ishape, ncoils, "i"), ishape, ncoils, False)
    spec = _shift(spec, ishape, ncoils, "f")
    crop = centered_crop(spec, oshape)
    del spec
    img = _fft3(_shift(crop, oshape, ncoils, "i"), oshape, ncoils, True)
    return _shift(img, oshape, ncoils, "f")


def resize_mask(A, oshape):
    """SigPy ``resize`` of the (nshots, N0, N1, 1) bool masks (pyramid.py:67)."""
    A = np.asarray(A)
    out = np.zeros(oshape, dtype=A.dtype)
    src, dst = [], []
    for i, o in zip(A.shape, oshape):
        ishift, oshift = max(i // 2 - o // 2, 0), max(o // 2 - i // 2, 0)
        n = min(i - ishift, o - oshift)
        src.append(slice(ishift, ishift + n))
        dst.append(slice(oshift, oshift + n))
    out[tuple(dst)] = A[tuple(src)]
    return out


def pyramid_reconstruction(ksp, mps, A, nshots, n_spatial_levels=3, n_temporal_levels=3, n_joint_iters=100,
                           save_path=None, sampling="disorder", final_cg_iter=100, compute_objective=True,
                           final_joint_iters=0, progress=None, tune=False):
    """pyramid.py:24-118.  Returns (x centred numpy image, t_est numpy
    (nshots, 6)).  ``n_temporal_levels`` is unused, as in the reference.
    Convergence histories are written to ``save_path`` as
    loss_<f>_factor.npy when it is given (the reference plots them).

    ``final_joint_iters`` > 0 replaces the reference's image-only final level
    (pyramid.py:96-115) by a JointRecon of that many outer iterations at full
    resolution starting from the coarse levels' motion (an extension: the
    reference never estimates motion at full resolution); its image and
    motion are returned.  ``pyramid_reconstruction.levels`` records each
    level's factor, shape, objective history, transform history and wall
    time; ``progress`` (a callable taking a message) is told when each level
    ends.  ``tune=True`` runs the autotuners (line-pass lines per workgroup,
    forward-encoding variant) on each level's geometry before its iterations;
    results do not depend on the choice, only the time does."""
    if tune:
        from .batched import tune_encoders, tune_problem

        def _tune(pb):
            tune_problem(pb)
            tune_encoders(pb)
    else:
        def _tune(pb):
            pass

    dev = _lib.require_gpu()
    ksp = as_c64(ksp, dev)
    mps = as_c64(mps, dev)
    A = A.cpu().numpy() if isinstance(A, torch.Tensor) else np.asarray(A)
    full = tuple(ksp.shape[1:])
    ncoils = ksp.shape[0]
    t_est = np.zeros((nshots, 6), dtype=float)
    levels = []
    import time

    for sl in range(n_spatial_levels - 1):
        t_level = time.perf_counter()
        f = 2 ** (n_spatial_levels - sl - 1)
        if sampling == "disorder":
            ds = (int(full[0] // f), int(full[1] // f), full[2])
        elif sampling == "sequential":
            ds = (int(full[0] // f), full[1], full[2])
        else:
            raise ValueError("sampling must be 'disorder' or 'sequential'")
        mps_res = resample_maps(mps, ds)
        ksp_res = centered_crop(ksp, ds)
        A_res = resize_mask(A, (nshots, *ds[:2], 1))
        kgrid, rkgrid = compute_transform_grids_voxel(full, [1, 1, 1], ds)
        app = JointRecon(ksp_res, mps_res, A_res, kgrid, rkgrid, t0=t_est, max_joint_iter=int(n_joint_iters * f),
                         compute_objective=compute_objective)
        _tune(app.problem)
        _, t = app.run()
        t_est = t.detach().cpu().numpy().copy()
        levels.append(dict(factor=f, shape=ds, objective=app.objective_history, transforms=app.transform_history,
                           seconds=time.perf_counter() - t_level))
        if progress is not None:
            progress(f"level x{f} {ds}: {int(n_joint_iters * f)} outer iterations in {levels[-1]['seconds']:.1f}s")
        if save_path is not None:
            os.makedirs(save_path, exist_ok=True)
            np.save(os.path.join(str(save_path), f"loss_{f}_factor.npy"), app.objective_history)
        del app, mps_res, ksp_res
    t_level = time.perf_counter()
    kgrid, rkgrid = compute_transform_grids_voxel(full, [1, 1, 1], None)
    if final_joint_iters and final_joint_iters > 0:
        app = JointRecon(ksp, mps, A, kgrid, rkgrid, t0=t_est, max_joint_iter=int(final_joint_iters),
                         compute_objective=compute_objective)
        _tune(app.problem)
        img, t = app.run()
        t_est = t.detach().cpu().numpy().copy()
        torch.cuda.synchronize(dev)
        levels.append(dict(factor=1, shape=full, objective=app.objective_history, transforms=app.transform_history,
                           seconds=time.perf_counter() - t_level))
        pyramid_reconstruction.levels = levels
        return img.cpu().numpy(), t_est
    # full resolution: image estimation only (pyramid.py:96-115)
    prob = SenseProblem(ksp, mps, A, kgrid, rkgrid, centered=True, normalize=True)
    _tune(prob)
    prob.theta.copy_(torch.as_tensor(t_est, device=dev))
    x = prob.zeros_image()
    prob.cg(x, None, int(final_cg_iter), 1e-6, check_every=1)  # b - A x0 in one normal operator
    img = prob.image_from_plane(x, shift="f", scale=prob.max_norm)
    torch.cuda.synchronize(dev)
    levels.append(dict(factor=1, shape=full, objective=None, transforms=None, seconds=time.perf_counter() - t_level))
    pyramid_reconstruction.levels = levels
    return img.cpu().numpy(), t_est


def aggregate_sampling_masks(A, bin_factor):
    """pyramid.py:120-123: merge consecutive shots in bins of ``bin_factor``
    by summing their masks (an integer count per line, as the reference's
    ``sum(1)``).  Used only by the reference's disabled temporal levels
    (pyramid.py:80-95); kept for API parity."""
    nshots = A.shape[0]
    nbins = nshots // bin_factor
    if nbins * bin_factor != nshots:  # the reference's reshape fails too
        raise ValueError(f"{nshots} shots do not split into bins of {bin_factor}")
    if isinstance(A, torch.Tensor):
        return A.reshape(nbins, bin_factor, *A.shape[1:]).sum(1)
    A = np.asarray(A)
    return A.reshape((nbins, bin_factor, *A.shape[1:])).sum(1)


def aggregate_transforms(T, bin_factor):
    """pyramid.py:125-128: per-bin SUM (not mean) of consecutive shots'
    parameters, as the reference computes it."""
    nshots = T.shape[0]
    nbins = nshots // bin_factor
    if nbins * bin_factor != nshots:
        raise ValueError(f"{nshots} shots do not split into bins of {bin_factor}")
    if isinstance(T, torch.Tensor):
        return T.reshape(nbins, bin_factor, 6).sum(1)
    T = np.asarray(T)
    return T.reshape((nbins, bin_factor, 6)).sum(1)
