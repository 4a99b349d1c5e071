"""Uncorrected SENSE reconstruction: the baseline the reference compares
against (simulation.py:462-466, recon.py:153), i.e. ``sigpy.mri.app.SenseRecon``
on the GPU.

SigPy solves  min_x 0.5 ||W^1/2 (F_c S x - y)||^2 + 0.5 lamda ||x||^2  with
its ConjugateGradient on (A^H A + lamda I) x = A^H W^1/2 y, A = W^1/2 F_c S
(F_c the centred orthonormal 3-D FFT), optional preconditioner P, x0 = 0 and
the weights, when not given, estimated as rss(y over coils) > 0.  Here the
weights must be binary and constant along the readout axis 2 (true for the
reference's data: whole readout lines are sampled); the sampled lines are
split into pseudo-shots of at most 512 lines (assigned round-robin so each
spans all columns) and the normal operator is the identity-motion encoding
E^H E of the motion-corrected path, summed over the pseudo-shots.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .encoding import CG_DONE, SenseProblem, real_to_plane
from .utils import _to_numpy, as_c64, compute_transform_grids_voxel

LINES_PER_GROUP = 512


def _line_groups(line_mask):
    """(N0, N1) bool -> (G, N0, N1, 1) disjoint masks of <= 512 lines each,
    lines dealt round-robin."""
    k0, k1 = np.nonzero(line_mask)
    n = k0.size
    if n == 0:
        raise ValueError("no sampled k-space lines")
    g = (n + LINES_PER_GROUP - 1) // LINES_PER_GROUP
    masks = np.zeros((g, *line_mask.shape, 1), dtype=bool)
    idx = np.arange(n) % g
    masks[idx, k0, k1, 0] = True
    return masks


def _weights_lines(ksp_d, weights, shape):
    """Binary readout-constant sampling weights -> (N0, N1) line mask
    (sigpy.mri.app._estimate_weights: rss(y) > 0 when weights is None)."""
    if weights is None:
        C = ksp_d.shape[0]
        vol = int(np.prod(shape))
        P = torch.empty(vol, dtype=torch.float32, device=ksp_d.device)
        M = torch.empty(vol, dtype=torch.uint8, device=ksp_d.device)
        # sum over coils of |y|^2 > 0 (the threshold of me_coil_ssq is strict)
        _lib.call("me_coil_ssq", _lib.ptr(ksp_d), C, vol, 1.0, 0.0, _lib.ptr(P), _lib.ptr(M), _lib.stream(ksp_d.device))
        w = M.cpu().numpy().reshape(shape).astype(bool)
    else:
        w = np.asarray(_to_numpy(weights))
        if w.shape != tuple(shape):
            w = np.broadcast_to(w, shape)
        if not np.all((w == 0) | (w == 1)):
            raise NotImplementedError("only binary sampling weights are supported")
        w = w.astype(bool)
    line = w.any(axis=2)
    if not np.array_equal(w, np.broadcast_to(line[..., None], w.shape)):
        raise NotImplementedError("sampling weights must be constant along the readout axis 2")
    return line


class SenseRecon:
    """sigpy.mri.app.SenseRecon(y, mps, lamda=0, weights=None, P=None,
    max_iter=100, tol=0).run() on the GPU.

    y, mps: (C, N0, N1, N2) centred k-space and maps (NumPy, torch or
    CuPy-like).  P: image-shaped real preconditioner (the diagonal of the
    reference's ``sp.linop.Multiply``) or None.  ``run()`` returns the image
    (N0, N1, N2) complex64 torch tensor in the frame of the maps."""

    def __init__(self, y, mps, lamda=0.0, weights=None, P=None, max_iter=100, tol=0.0, x0=None, show_pbar=True,
                 **kwargs):
        dev = _lib.require_gpu()
        ksp_d = as_c64(y, dev)
        shape = tuple(ksp_d.shape[1:])
        self.shape = shape
        self.lamda = float(lamda)
        self.max_iter = int(max_iter)
        self.tol = float(tol)
        # pseudo-shot masks in the centred frame (SenseProblem uncentres
        # data, maps and masks together)
        line_c = _weights_lines(ksp_d, weights, shape)
        masks = _line_groups(line_c)
        kg, rkg = compute_transform_grids_voxel(shape, (1, 1, 1))
        ones = np.ones(shape, dtype=np.float32)
        self.pb = SenseProblem(ksp_d, mps, masks, kg, rkg, P=ones, M=ones, centered=True, normalize=False)
        del ksp_d
        pb = self.pb
        if P is not None:  # centred frame -> uncentred plane-major
            p = np.fft.ifftshift(np.asarray(_to_numpy(P), dtype=np.float32).reshape(shape))
            self.P = real_to_plane(p, shape, torch.float32, dev)
        else:
            self.P = None
        self.x = pb.zeros_image() if x0 is None else pb.image_to_plane(as_c64(x0, dev), shift="i")

    def _normal(self, p, out):
        """out = sum_g E_g^H E_g p + lamda p (identity motion, no output mask)."""
        pb = self.pb
        U = pb._ubuf()
        Y = pb._ybuf()
        s = _lib.stream(pb.device)
        for s0, s1 in pb.batches():
            pb._encode(p, 0, s0, s1, y=Y)
            pb._decode(Y, s0, s1, U)
            _lib.call("me_sum_batch", _lib.ptr(U), pb.vol * pb.ncg, (s1 - s0) * pb.ncg, pb.vol, _lib.ptr(out),
                      1 if s0 > 0 else 0, s)
        if self.lamda != 0.0:
            _lib.call("me_axpy_real", _lib.ptr(out), self.lamda, _lib.ptr(p), pb.vol, s)
        return out

    def _rhs(self):
        pb = self.pb
        U = pb._ubuf()
        b = pb.zeros_image()
        s = _lib.stream(pb.device)
        for s0, s1 in pb.batches():
            pb._decode(pb.ymeas, s0, s1, U)
            _lib.call("me_sum_batch", _lib.ptr(U), pb.vol * pb.ncg, (s1 - s0) * pb.ncg, pb.vol, _lib.ptr(b),
                      1 if s0 > 0 else 0, s)
        return b

    def run(self):
        pb = self.pb
        n = pb.vol
        s = _lib.stream(pb.device)
        st = pb._cg_state
        st.zero_()
        st[7] = 1.0  # sigpy.alg.ConjugateGradient stopping rule
        st[8] = self.tol
        b = self._rhs()
        x = self.x
        Ap = torch.empty_like(x)
        r = torch.empty_like(x)
        p = torch.empty_like(x)
        self._normal(x, Ap)
        _lib.call("me_cg_init", _lib.ptr(b), _lib.ptr(Ap), _lib.ptr(r), _lib.ptr(p), _lib.ptr(self.P), n, _lib.ptr(st),
                  _lib.ptr(pb._vec_partials), s)
        for it in range(self.max_iter):
            if float(st[CG_DONE].item()) != 0.0:
                break
            self._normal(p, Ap)
            _lib.call("me_cg_step", _lib.ptr(x), _lib.ptr(r), _lib.ptr(p), _lib.ptr(Ap), _lib.ptr(self.P), n,
                      self.tol, _lib.ptr(st), _lib.ptr(pb._vec_partials), None, None, 0, s)
        self.iterations = int(st[6].item())
        return pb.image_from_plane(x, shift="f")


def sense_recon(y, mps, lamda=0.0, weights=None, P=None, max_iter=100, tol=0.0, x0=None):
    """Functional form of :class:`SenseRecon`."""
    return SenseRecon(y, mps, lamda=lamda, weights=weights, P=P, max_iter=max_iter, tol=tol, x0=x0).run()
