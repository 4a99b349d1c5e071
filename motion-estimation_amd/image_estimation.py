"""CG-SENSE image estimation (image_estimation.py:10-165) on the GPU."""
from __future__ import annotations

from .encoding import SenseProblem
from .utils import as_f64


def estimate_image_cg(ksp, mps, shot_mask, transforms, kgrid, rkgrid, P, M, x0=None, max_iter=100, tol=1e-6,
                      show_pbar=True):
    """Reconstruct an image by preconditioned CG on A x = b
    (image_estimation.py:10-165).

    Same arguments as the reference: inputs in the uncentred (ifftshifted)
    frame; ksp/mps (ncoils, N0, N1, N2); shot_mask (nshots, N0, N1, 1);
    transforms (nshots, 6); P, M image-shaped.  A = M * sum_s T_s^H E_s^H E_s
    T_s, b = M * sum_s T_s^H E_s^H y_s, z = P r, stop when sqrt|<r, z>| < tol
    (checked after the x/r update, as the reference).  Returns the image as a
    complex64 torch tensor (N0, N1, N2) on the GPU.  ``show_pbar`` is accepted
    for signature compatibility.
    """
    prob = SenseProblem(ksp, mps, shot_mask, kgrid, rkgrid, P=P, M=M, centered=False, normalize=False)
    prob.theta.copy_(as_f64(transforms, prob.device).reshape(prob.nshots, 6))
    x = prob.zeros_image() if x0 is None else prob.image_to_plane(x0)
    prob.cg(x, None, int(max_iter), float(tol))  # b - A x0 in one normal operator
    return prob.image_from_plane(x)
