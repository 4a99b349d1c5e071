"""ESPIRiT receive-sensitivity maps (espirit.py:47-239) on the GPU.

The reference takes the SVD of the calibration matrix H, keeps the left
singular vectors above ``eig_thresh`` times the largest singular value,
zero-pads them to the image size, inverse-FFTs them and takes each voxel's
principal coil eigenvector.  Here the left singular subspace comes from a
Jacobi eigendecomposition of G = H H^H (eigenvalues = squared singular
values) and the maps depend on it only through the projector P = U_r U_r^H,
so the choice of basis inside the kept subspace (LAPACK's or Jacobi's) does
not change the result.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .utils import as_c64

JACOBI_TOL = 1e-12


def _ints(vals):
    return (C.c_int * len(vals))(*[int(v) for v in vals])


def rx_espirit_3d(calib, imsize, kernel_size=None, eig_thresh=0.02, mask_thresh=0.99, max_sweeps=40,
                  power_iters=2000):
    """espirit.py:47-239.  calib: (NKx, NKy, NKz, NRx) complex k-space
    (NumPy, torch or CuPy-like); returns the maps (Nx, Ny, Nz, NRx) as a
    complex64 torch tensor on the current HIP device.

    ``eig_thresh`` < 1 keeps singular values s > s[0] * eig_thresh, otherwise
    the round(eig_thresh) largest (at most min(rows, blocks), the columns of
    the reference's economy SVD).  With nothing kept the maps are zero, as
    in the reference.  After a call, ``rx_espirit_3d.singular_values``
    (descending), ``.kept`` and ``.sweeps`` describe the decomposition."""
    dev = _lib.require_gpu()
    if kernel_size is None:
        kernel_size = (5, 5, 5)
    kernel = tuple(int(k) for k in kernel_size)
    cal = as_c64(calib, dev).contiguous()
    if cal.dim() != 4:
        raise ValueError("calib must be (NKx, NKy, NKz, NRx)")
    NRx = int(cal.shape[3])
    imsize = tuple(int(v) for v in imsize)
    dims = _ints(cal.shape)
    kern = _ints(kernel)
    s = _lib.stream(dev)
    ncol = int(_lib.size("me_espirit_hankel_size", dims, kern))
    if ncol <= 0:
        raise ValueError(f"kernel {kernel} does not fit the calibration region {tuple(cal.shape[:3])}")
    m = int(np.prod(kernel)) * NRx
    nb = ncol // m
    H = torch.empty(ncol, dtype=torch.complex64, device=dev)
    _lib.call("me_espirit_calibration_matrix", _lib.ptr(cal), dims, kern, _lib.ptr(H), s)
    G = torch.empty((m, m, 2), dtype=torch.float64, device=dev)
    _lib.call("me_gram_rows", _lib.ptr(H), m, nb, _lib.ptr(G), s)
    del H
    V = torch.empty((m, m, 2), dtype=torch.float64, device=dev)
    lam = torch.empty(m, dtype=torch.float64, device=dev)
    work = torch.empty(int(_lib.size("me_jacobi_workspace_size", m)), dtype=torch.float64, device=dev)
    sweeps = C.c_int(0)
    _lib.call("me_jacobi_eig", _lib.ptr(G), _lib.ptr(V), m, int(max_sweeps), JACOBI_TOL, _lib.ptr(lam),
              _lib.ptr(work), C.byref(sweeps), s)
    del G, work
    lam_h = lam.cpu().numpy()
    order = np.argsort(-lam_h, kind="stable")[:min(m, nb)]
    sv = np.sqrt(np.maximum(lam_h[order], 0.0))
    if eig_thresh < 1:
        keep = order[sv > sv[0] * eig_thresh] if sv.size and sv[0] > 0 else order[:0]
    else:
        keep = order[:int(round(eig_thresh))]
    rx_espirit_3d.singular_values = sv
    rx_espirit_3d.kept = int(keep.size)
    rx_espirit_3d.sweeps = int(sweeps.value)
    sens = torch.zeros((*imsize, NRx), dtype=torch.complex64, device=dev)
    if keep.size == 0:
        return sens
    r = int(keep.size)
    idx = torch.as_tensor(keep.astype(np.int32), device=dev)
    X = torch.empty(m * r, dtype=torch.complex64, device=dev)
    P = torch.empty((m, m, 2), dtype=torch.float64, device=dev)
    _lib.call("me_espirit_projector", _lib.ptr(V), m, _lib.ptr(idx), r, _lib.ptr(X), _lib.ptr(P), s)
    del V, X
    isz = _ints(imsize)
    mw = torch.empty(int(_lib.size("me_espirit_maps_workspace_size", NRx, kern, isz)), dtype=torch.complex64,
                     device=dev)
    _lib.call("me_espirit_maps", _lib.ptr(P), NRx, kern, isz, float(mask_thresh), int(power_iters), _lib.ptr(mw),
              _lib.ptr(sens), s)
    return sens
