"""Rigid-transform operators over the HIP line-pass kernels.

Drop-in counterparts of the reference's non-legacy operators:

* ``RigidTransformCudaOptimzied`` (transform.py:243-377): ``.apply(x, out=None)``
  and ``.adjoint(x, out=None)`` returning new arrays;
* ``RigidTransformDerivativeCuda`` (transform.py:379-504): ``.apply(input, p_idx)``.

Arrays may be torch tensors, NumPy or CuPy-like arrays; results are complex64
torch tensors on the current HIP device.  Every transform runs on the GPU
through libmeamd.so; there is no CPU path.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .utils import as_c64, as_f64, grid_vectors


class Geometry:
    """Level shape, device grid vectors and strides of one volume layout."""

    _cache = {}

    def __init__(self, shape, rvec, device):
        self.shape = tuple(int(s) for s in shape)
        self.device = device
        self.r = [torch.as_tensor(np.ascontiguousarray(v, dtype=np.float64), device=device) for v in rvec]
        n0, n1, n2 = self.shape
        self.st_std = (n1 * n2, n2, 1)  # reference C order (x, y, z)
        self.st_plane = (n1, 1, n0 * n1)  # plane-major: contiguous (x, y) planes per z
        self.numel = n0 * n1 * n2

    @classmethod
    def from_grids(cls, kgrid, rkgrid, device=None):
        device = device or _lib.require_gpu()
        shape, rvec = grid_vectors(kgrid, rkgrid)
        key = (shape, tuple(np.asarray(v).tobytes() for v in rvec), str(device))
        g = cls._cache.get(key)
        if g is None:
            if len(cls._cache) > 64:
                cls._cache.clear()
            g = cls._cache[key] = cls(shape, rvec, device)
        return g

    def n_arg(self):
        return _lib.int3(self.shape)

    def st_arg(self, plane=False):
        return _lib.ll3(self.st_plane if plane else self.st_std)

    def r_args(self):
        return [_lib.ptr(v) for v in self.r]


def rigid_batch(x, theta, geom, mode=0, out=None, plane=False, sum_batch=False, mask=None, work=None):
    """Batched T (mode 0), T^H (mode 1) or T with inverse translation (mode 2).

    x: (B, N0, N1, N2) complex64 device tensor (or (N0,N1,N2) with B=1) in the
    reference layout, or plane-major layout if ``plane``.  theta: (B, 6)
    float64 device tensor.  With ``sum_batch`` the adjoint outputs are summed
    (and multiplied by ``mask``) into one volume."""
    B = theta.shape[0]
    vol = geom.numel
    if out is None:
        out = torch.empty((vol,) if sum_batch else (B, vol), dtype=torch.complex64, device=x.device)
    if sum_batch and work is None:
        work = torch.empty((B, vol), dtype=torch.complex64, device=x.device)
    coef = torch.empty((B, 16), dtype=torch.float64, device=x.device)
    flags = (1 if sum_batch else 0) | (2 if mask is not None else 0)
    _lib.call("me_rigid_transform", _lib.ptr(x), vol if x.numel() > vol else 0, 1, _lib.ptr(out), vol,
              _lib.ptr(work), vol, int(B), geom.n_arg(), geom.st_arg(plane), *geom.r_args(), _lib.ptr(theta),
              int(theta.stride(0)), _lib.ptr(coef), int(mode), int(flags), _lib.ptr(mask), _lib.stream(x.device))
    return out


def rigid_deriv_batch(x, theta, geom, p_idx, out=None, work=None, plane=False):
    B = theta.shape[0]
    vol = geom.numel
    if out is None:
        out = torch.empty((B, vol), dtype=torch.complex64, device=x.device)
    if work is None and p_idx >= 3:
        work = torch.empty((B, vol), dtype=torch.complex64, device=x.device)
    coef = torch.empty((B, 16), dtype=torch.float64, device=x.device)
    _lib.call("me_rigid_deriv", _lib.ptr(x), vol if x.numel() > vol else 0, _lib.ptr(out), vol, _lib.ptr(work),
              vol, int(B), geom.n_arg(), geom.st_arg(plane), *geom.r_args(), _lib.ptr(theta), int(theta.stride(0)),
              _lib.ptr(coef), int(p_idx), _lib.stream(x.device))
    return out


def _params6(parameters, device):
    t = as_f64(parameters, device).reshape(-1)
    if t.numel() != 6:
        raise ValueError("parameters must have 6 entries [q1, q2, q3, t1, t2, t3]")
    return t.reshape(1, 6)


def _check_shape(x, geom):
    if tuple(x.shape) != geom.shape:
        raise ValueError(f"volume shape {tuple(x.shape)} does not match the grid shape {geom.shape}")


class RigidTransformCudaOptimzied:
    """transform.py:243-377.  T(theta) x = Tr(q) Rz Ry Rx x with FFT shears.

    ``inverse=True`` conjugates only the translation in ``apply``
    (transform.py:308); ``adjoint`` always conjugates (transform.py:332)."""

    def __init__(self, parameters, kgrid, rkgrid, inverse=False):
        self.params = parameters
        self.kgrid = kgrid
        self.rkgrid = rkgrid
        self.inverse = inverse
        self._geom = None

    def _g(self):
        if self._geom is None:
            self._geom = Geometry.from_grids(self.kgrid, self.rkgrid)
        return self._geom

    def _run(self, x, out, mode):
        g = self._g()
        src = as_c64(out if out is not None else x, g.device)  # quirk Q2: transforms out's contents
        _check_shape(src, g)
        th = _params6(self.params, g.device)
        y = rigid_batch(src, th, g, mode=mode)
        return y.reshape(g.shape)

    def apply(self, x, out=None):
        """transform.py:263-318 (returns a new array)."""
        return self._run(x, out, 2 if self.inverse else 0)

    def adjoint(self, x, out=None):
        """transform.py:320-377 (returns a new array)."""
        return self._run(x, out, 1)


class RigidTransformDerivativeCuda:
    """transform.py:379-504.  ``apply(input, p_idx)`` returns
    d/d(theta_p) [Tr Rx Ry Rz] input (rotations z, y, x, the reverse of
    RigidTransformCudaOptimzied.apply).  p_idx outside [0, 5] returns a copy of
    the input, as the reference's missing else branch does."""

    def __init__(self, shape, parameters, kgrid, rkgrid):
        self.parameters = parameters
        self.kgrid = kgrid
        self.rkgrid = rkgrid
        self.shape = tuple(shape)
        self._geom = Geometry.from_grids(kgrid, rkgrid)
        g = self._geom
        # preallocated output and work volume (transform.py:388-390)
        self._work = torch.empty((1, g.numel), dtype=torch.complex64, device=g.device)

    def apply(self, input, p_idx):
        g = self._geom
        x = as_c64(input, g.device)
        _check_shape(x, g)
        if not (0 <= int(p_idx) <= 5):
            return x.clone()
        th = _params6(self.parameters, g.device)
        y = rigid_deriv_batch(x, th, g, int(p_idx), work=self._work)
        return y.reshape(g.shape)
