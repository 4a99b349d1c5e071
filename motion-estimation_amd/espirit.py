"""ESPIRiT receive-sensitivity maps (espirit.py:47-239) on the GPU.

The reference takes the SVD of the calibration matrix H, keeps the left
singular vectors above ``eig_thresh`` times the largest singular value,
zero-pads them to the image size, inverse-FFTs them and takes each voxel's
principal coil eigenvector.  Here the left singular subspace is that of
G = H H^H (eigenvalues = squared singular values), and the maps depend on
it only through the projector P = U_r U_r^H, so the choice of basis inside
the kept subspace (LAPACK's or this one) does not change the result.

Two ways to the kept subspace:

* ``"jacobi"``: all of G by parallel two-sided Jacobi (m x m, 20 or so
  sweeps of m - 1 rounds).
* ``"subspace"``: block subspace iteration with Rayleigh-Ritz.  A random
  m x k block is multiplied by G three times, orthonormalised after each
  product (Loewdin: Y (V D^-1/2) from the eigendecomposition of Y^H Y), and
  Q^H G Q (k x k) is diagonalised by Jacobi.  Accepted when the smallest
  Ritz value is below 1e-3 of the keep threshold (the block then reaches
  well under the threshold, so every kept eigenvector lies in it to
  round-off: the error falls as (lambda_k / lambda_kept)^5), or when G's
  rank is below k; otherwise k doubles, and past m/2 the Jacobi route runs.
  ``"auto"`` uses it for m > 1024 with a ratio threshold.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .utils import as_c64

JACOBI_TOL = 1e-12
SUBSPACE_MIN_ROWS = 1024
SUBSPACE_BLOCK = 512
SUBSPACE_POWER = 2
RANK_CUTOFF = 1e-14  # relative Gram eigenvalue below which a block direction is dropped


def _ints(vals):
    return (C.c_int * len(vals))(*[int(v) for v in vals])


class _Ops:
    """FP64 complex matrices as (rows, cols, 2) float64 device tensors."""

    def __init__(self, dev):
        self.dev = dev
        self.s = _lib.stream(dev)

    def gemm(self, opa, opb, A, B, m, n, k):
        Cm = torch.empty((m, n, 2), dtype=torch.float64, device=self.dev)
        _lib.call("me_zgemm", opa, opb, m, n, k, _lib.ptr(A), _lib.ptr(B), _lib.ptr(Cm), self.s)
        return Cm

    def eig(self, A, n, inplace=False, max_sweeps=40):
        """Hermitian eigendecomposition (Jacobi): eigenvalues (host, unsorted)
        and eigenvectors (columns, device)."""
        G = A if inplace else A.clone()
        V = torch.empty((n, n, 2), dtype=torch.float64, device=self.dev)
        lam = torch.empty(n, dtype=torch.float64, device=self.dev)
        work = torch.empty(int(_lib.size("me_jacobi_workspace_size", n)), dtype=torch.float64, device=self.dev)
        sweeps = C.c_int(0)
        _lib.call("me_jacobi_eig", _lib.ptr(G), _lib.ptr(V), n, int(max_sweeps), JACOBI_TOL, _lib.ptr(lam),
                  _lib.ptr(work), C.byref(sweeps), self.s)
        return lam.cpu().numpy(), V, int(sweeps.value)

    def orth(self, Y, m, k):
        """Orthonormal basis of Y's columns (Loewdin); dropped directions become
        zero columns.  Returns (Q, rank)."""
        S = self.gemm(1, 0, Y, Y, k, k, m)
        lam, V, _ = self.eig(S, k, inplace=True)
        top = max(float(lam.max()), 0.0)
        keep = lam > RANK_CUTOFF * top if top > 0 else np.zeros(k, dtype=bool)
        scale = np.where(keep, 1.0 / np.sqrt(np.where(keep, lam, 1.0)), 0.0)
        sd = torch.as_tensor(scale, device=self.dev)
        _lib.call("me_scale_columns", _lib.ptr(V), k, k, _lib.ptr(sd), self.s)
        return self.gemm(0, 0, Y, V, m, k, k), int(keep.sum())


def _keep(lam_sorted, order, eig_thresh, limit):
    sv = np.sqrt(np.maximum(lam_sorted, 0.0))
    if eig_thresh < 1:
        keep = order[sv > sv[0] * eig_thresh] if sv.size and sv[0] > 0 else order[:0]
    else:
        keep = order[:min(int(round(eig_thresh)), limit)]
    return sv, keep


def _subspace(ops, G, m, nb, eig_thresh, seed=20240601):
    """Kept eigenpairs of G by subspace iteration, or None when the block
    cannot be certified (then the caller runs full Jacobi)."""
    k = min(SUBSPACE_BLOCK, m)
    while 2 * k <= m:
        Om = torch.empty((m, k, 2), dtype=torch.float64, device=ops.dev)
        _lib.call("me_random_normal", _lib.ptr(Om), m * k, C.c_ulonglong(seed + k), ops.s)
        Y = ops.gemm(0, 0, G, Om, m, k, m)
        del Om
        for _ in range(SUBSPACE_POWER):
            Q, _ = ops.orth(Y, m, k)
            Y = ops.gemm(0, 0, G, Q, m, k, m)
        Q, rank = ops.orth(Y, m, k)
        del Y
        B = ops.gemm(1, 0, Q, ops.gemm(0, 0, G, Q, m, k, m), k, k, m)
        lam, W, sweeps = ops.eig(B, k, inplace=True)
        order = np.argsort(-lam, kind="stable")[:min(rank, nb)]
        sv, keep = _keep(lam[order], order, eig_thresh, rank)
        ritz_min = float(lam[order[-1]]) if order.size else 0.0
        tau = (eig_thresh ** 2) * float(lam[order[0]]) if eig_thresh < 1 and order.size else None
        enough = rank < k or (tau is not None and ritz_min < 1e-3 * tau) or \
            (eig_thresh >= 1 and 0 < keep.size and keep.size + 64 <= rank
             and ritz_min < 1e-3 * float(lam[keep[-1]]))
        if enough:
            return ops.gemm(0, 0, Q, W, m, k, k), k, sv, keep, sweeps
        k *= 2
    return None


def rx_espirit_3d(calib, imsize, kernel_size=None, eig_thresh=0.02, mask_thresh=0.99, max_sweeps=40,
                  power_iters=2000, method="auto"):
    """espirit.py:47-239.  calib: (NKx, NKy, NKz, NRx) complex k-space
    (NumPy, torch or CuPy-like); returns the maps (Nx, Ny, Nz, NRx) as a
    complex64 torch tensor on the current HIP device.

    ``eig_thresh`` < 1 keeps singular values s > s[0] * eig_thresh, otherwise
    the round(eig_thresh) largest (at most min(rows, blocks), the columns of
    the reference's economy SVD).  With nothing kept the maps are zero, as
    in the reference.  ``method``: "auto", "jacobi" or "subspace" (module
    docstring).  After a call, ``rx_espirit_3d.singular_values`` (descending;
    with the subspace route only the block's Ritz values), ``.kept``,
    ``.sweeps`` and ``.method`` describe the decomposition."""
    dev = _lib.require_gpu()
    if method not in ("auto", "jacobi", "subspace"):
        raise ValueError(f"unknown method {method!r}")
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
    ops = _Ops(dev)
    s = ops.s
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
    res = None
    use_sub = method == "subspace" or (method == "auto" and m > SUBSPACE_MIN_ROWS and eig_thresh < 1)
    if use_sub:
        res = _subspace(ops, G, m, nb, eig_thresh)
    if res is not None:
        U, ldu, sv, keep, sweeps = res
        used = "subspace"
    else:
        lam, U, sweeps = ops.eig(G, m, inplace=True, max_sweeps=max_sweeps)
        ldu = m
        order = np.argsort(-lam, kind="stable")[:min(m, nb)]
        sv, keep = _keep(lam[order], order, eig_thresh, min(m, nb))
        used = "jacobi"
    del G
    rx_espirit_3d.singular_values = sv
    rx_espirit_3d.kept = int(keep.size)
    rx_espirit_3d.sweeps = sweeps
    rx_espirit_3d.method = used
    sens = torch.zeros((*imsize, NRx), dtype=torch.complex64, device=dev)
    if keep.size == 0:
        return sens
    r = int(keep.size)
    idx = torch.as_tensor(keep.astype(np.int32), device=dev)
    X = torch.empty(m * r, dtype=torch.complex64, device=dev)
    P = torch.empty((m, m, 2), dtype=torch.float64, device=dev)
    _lib.call("me_espirit_projector", _lib.ptr(U), m, ldu, _lib.ptr(idx), r, _lib.ptr(X), _lib.ptr(P), s)
    del U, X
    isz = _ints(imsize)
    mw = torch.empty(int(_lib.size("me_espirit_maps_workspace_size", NRx, kern, isz)), dtype=torch.complex64,
                     device=dev)
    _lib.call("me_espirit_maps", _lib.ptr(P), NRx, kern, isz, float(mask_thresh), int(power_iters), _lib.ptr(mw),
              _lib.ptr(sens), s)
    return sens
