"""Levenberg-Marquardt motion estimation (transform_estimation.py:7-152) on the GPU."""
from __future__ import annotations

import numpy as np
import torch

from .encoding import SenseProblem
from .utils import as_f64


def _write_back(dst, src):
    """Copy a device tensor into the caller's array in place (torch, NumPy
    or CuPy-like), as the reference mutates its arguments."""
    if isinstance(dst, torch.Tensor):
        dst.copy_(src.to(dst.device, dst.dtype).reshape(dst.shape))
    else:
        val = src.detach().cpu().numpy().reshape(np.shape(dst))
        if hasattr(dst, "set"):  # CuPy-like
            dst.set(val.astype(dst.dtype))
        else:
            dst[...] = val


def estimate_transform(ksp, mps, shot_mask, image, kgrid, rkgrid, damp, convergence_flags, residual_buf=None,
                       partial_buf=None, gradient_buf=None, hessian_buf=None, transforms=None, max_iter=10, tol=1e-6,
                       use_ga=False):
    """One LM update of every shot's rigid parameters
    (transform_estimation.py:7-152), then mean re-registration.

    Mutates ``transforms``, ``damp`` and ``convergence_flags`` in place and
    returns T(mean) image (complex64 torch tensor on the GPU).  The dense
    ``residual_buf``/``partial_buf``/``gradient_buf``/``hessian_buf`` of the
    reference are accepted and not needed: residuals and Jacobian columns are
    kept only at the sampled lines.  ``max_iter`` is unused, as in the
    reference (Q9).  Shots are updated in parallel, which equals the
    reference's sequential loop because the image is fixed during it.
    ``use_ga`` enables the reference's experimental geodesic acceleration
    (transform_estimation.py:73-121).
    """
    if transforms is None:  # transform_estimation.py:41 indexes transforms[s]
        raise TypeError("'NoneType' object is not subscriptable: transforms is required")
    prob = SenseProblem(ksp, mps, shot_mask, kgrid, rkgrid, centered=False, normalize=False)
    prob.theta.copy_(as_f64(transforms, prob.device).reshape(prob.nshots, 6))
    prob.damp.copy_(as_f64(damp, prob.device).reshape(prob.nshots))
    fl = torch.as_tensor(np.asarray(convergence_flags if not isinstance(convergence_flags, torch.Tensor)
                                    else convergence_flags.cpu().numpy()).astype(np.uint8))
    prob.flags.copy_(fl.to(prob.device))
    x = prob.image_to_plane(image)
    out = prob.lm_step(x, tol=tol, use_ga=bool(use_ga))
    _write_back(transforms, prob.theta)
    _write_back(damp, prob.damp)
    _write_back(convergence_flags, prob.flags.bool())
    return prob.image_from_plane(out)
