"""Calibration preprocessing on the GPU: noise whitening (noise_whiten.py:5-14)
and SVD coil compression (coil_compression.py:6-46).

Both work on coil-major arrays (ncoils, ...) and return complex64 torch
tensors on the current HIP device; the C x C matrices are FP64."""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .utils import as_c64


def _gram(x, ncoils, n, w=None, scale=1.0):
    dev = x.device
    part = torch.empty(int(_lib.size("me_coil_gram_partials_size", ncoils)), dtype=torch.float64, device=dev)
    G = torch.empty((ncoils, ncoils, 2), dtype=torch.float64, device=dev)
    _lib.call("me_coil_gram", _lib.ptr(x), ncoils, n, _lib.ptr(w), float(scale), _lib.ptr(part), _lib.ptr(G),
              _lib.stream(dev))
    return G


def _mix(M, x, cout, cin, n):
    """y = M x over coils: M (cout, cin) complex, x (cin, n)."""
    y = torch.empty((cout, n), dtype=torch.complex64, device=x.device)
    Mc = M.to(torch.complex64).contiguous()
    _lib.call("me_coil_mix", _lib.ptr(Mc), cout, cin, _lib.ptr(x), n, _lib.ptr(y), _lib.stream(x.device))
    return y


def noise_whiten(ksp, noise, reg=1e-6):
    """noise_whiten.py:5-14: cov = noise noise^H / n + reg I, L = chol(cov),
    W = inv(L)^H, ksp_white = W ksp (the reference's W, applied as it is).
    Returns (ksp_white, W) with W an (ncoils, ncoils) complex128 tensor."""
    dev = _lib.require_gpu()
    k = as_c64(ksp, dev)
    nz = as_c64(noise, dev)
    C = k.shape[0]
    nn = nz.numel() // C
    nv = k.numel() // C
    G = _gram(nz.reshape(C, nn), C, nn, scale=1.0 / nn)
    W = torch.empty((C, C, 2), dtype=torch.float64, device=dev)
    _lib.call("me_whitening_matrix", _lib.ptr(G), C, float(reg), _lib.ptr(W), _lib.stream(dev))
    Wc = torch.view_as_complex(W)
    out = _mix(Wc, k.reshape(C, nv), C, C, nv).reshape(k.shape)
    return out, Wc


def coil_compression(mps, ksp=None, perc=0.95, reg=1e-6):
    """coil_compression.py:6-46: P = (S norm) S^H with norm = 1/(sum |S|^2 + reg),
    its SVD U S V^H (P is Hermitian PSD, so an eigendecomposition with
    eigenvalues descending), r = argmax(cumsum(S)/sum(S) >= perc) kept
    virtual coils (the reference's index, not a count + 1), A = U^H[:r],
    cc_mps = A S, cc_ksp = A ksp.  Virtual-coil phases follow the Jacobi
    eigenvectors; any unit phase per virtual coil gives the same
    reconstruction.  Returns (cc_mps, cc_ksp or None)."""
    dev = _lib.require_gpu()
    m = as_c64(mps, dev)
    C = m.shape[0]
    img = tuple(m.shape[1:])
    nv = m.numel() // C
    mf = m.reshape(C, nv)
    P = torch.empty(nv, dtype=torch.float32, device=dev)
    _lib.call("me_coil_ssq", _lib.ptr(mf), C, nv, float(reg), 0.0, _lib.ptr(P), None, _lib.stream(dev))
    G = _gram(mf, C, nv, w=P)
    lam = torch.empty(C, dtype=torch.float64, device=dev)
    U = torch.empty((C, C, 2), dtype=torch.float64, device=dev)
    _lib.call("me_hermitian_eig", _lib.ptr(G), C, 60, _lib.ptr(lam), _lib.ptr(U), _lib.stream(dev))
    lam_h = lam.cpu().numpy()
    energy = np.cumsum(lam_h) / np.sum(lam_h)
    r = int(np.argmax(energy >= perc))
    A = torch.view_as_complex(U).conj().T[:r].contiguous()
    if r == 0:  # the reference's index rule can keep no coil
        empty = torch.empty((0, *img), dtype=torch.complex64, device=dev)
        return empty, (None if ksp is None else empty)
    cc_mps = _mix(A, mf, r, C, nv).reshape(r, *img)
    cc_ksp = None
    if ksp is not None:
        k = as_c64(ksp, dev)
        cc_ksp = _mix(A, k.reshape(C, -1), r, C, k.numel() // C).reshape(r, *img)
    coil_compression.singular_values = lam_h
    coil_compression.matrix = A
    return cc_mps, cc_ksp
