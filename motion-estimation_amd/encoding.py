"""Device-resident SENSE problem: sample plan, encoding, CG and LM steps.

This is the engine behind the reference-named entry points
(``estimate_image_cg``, ``estimate_transform``, ``JointRecon``).  It keeps
every array on the GPU in the library's plane-major layout (each readout
plane x2 a contiguous N0 x N1 block), holds the measured data only at the
sampled (k0, k1) lines in hybrid (k0, k1, x2) space, and drives the HIP
kernels through the C ABI.  Shots are processed in batches sized to a memory
budget; within a batch every kernel launch covers all its shots.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .transform import Geometry
from .utils import as_c64, _to_numpy

CG_DONE = 5


# Working-set budget (bytes) that sets how many shots a batch holds: the
# per-shot work volumes (transformed images, partial images) of one batch
# stay within it.  8 GiB keeps all 64 shots of a 256 x 256 x 192 problem in
# one batch.
DEFAULT_MEM_BUDGET = 8 << 30

# Central-difference step (in units of v = -delta) of the geodesic
# acceleration's directional derivative J'v.
GA_STEP = 1e-3


def _i4(v):
    return (C.c_int * 4)(*[int(a) for a in v])


def _ll4(v):
    return (C.c_longlong * 4)(*[int(a) for a in v])


def std_strides(C_, shape):
    n0, n1, n2 = shape
    return (n0 * n1 * n2, n1 * n2, n2, 1)


def plane_strides(C_, shape):
    n0, n1, n2 = shape
    return (n0 * n1 * n2, n1, 1, n0 * n1)


def permute(x, xs, xd, y, ys, yd, off=(0, 0, 0, 0), scale=1.0):
    """y[c] = scale * x[(c + off) mod xd] (me_permute_shift)."""
    _lib.call("me_permute_shift", _lib.ptr(x), _ll4(xs), _i4(xd), _lib.ptr(y), _ll4(ys), _i4(yd), _i4(off),
              float(scale), _lib.stream(x.device))


def shift_offsets(shape, kind):
    """Per-axis offsets of ifftshift ('i') / fftshift ('f') / none (None)."""
    if kind is None:
        return (0, 0, 0, 0)
    if kind == "i":
        return (0, *[n // 2 for n in shape])
    return (0, *[n - n // 2 for n in shape])


def to_plane(x, shape, shift=None, scale=1.0, out=None, ncoils=1):
    """Reference C-order (C, N0, N1, N2) -> plane-major, optional (i)fftshift."""
    dims = (ncoils, *shape)
    if out is None:
        out = torch.empty(ncoils * int(np.prod(shape)), dtype=torch.complex64, device=x.device)
    permute(x, std_strides(ncoils, shape), dims, out, plane_strides(ncoils, shape), dims, shift_offsets(shape, shift),
            scale)
    return out


def from_plane(x, shape, shift=None, scale=1.0, ncoils=1):
    """Plane-major -> reference C-order, optional (i)fftshift."""
    dims = (ncoils, *shape)
    out = torch.empty((ncoils, *shape) if ncoils > 1 else shape, dtype=torch.complex64, device=x.device)
    permute(x, plane_strides(ncoils, shape), dims, out, std_strides(ncoils, shape), dims, shift_offsets(shape, shift),
            scale)
    return out


def real_to_plane(a, shape, dtype, device):
    """Real image-shaped array (N0, N1, N2) -> flat plane-major tensor."""
    t = torch.as_tensor(_to_numpy(a) if not isinstance(a, torch.Tensor) else a).to(device)
    t = t.reshape(shape).permute(2, 0, 1).contiguous().to(dtype)
    return t.reshape(-1)


class ShotPlan:
    """Sampled (k0, k1) lines of every shot from masks (S, N0, N1[, 1])
    (simulation.py:396-402, uncentred frame), sorted per shot by (k1, k0)."""

    def __init__(self, shot_mask, device):
        m = _to_numpy(shot_mask)
        if m.ndim == 4:
            if m.shape[3] != 1:
                if not np.all(m == m[..., :1]):
                    raise ValueError("shot masks must not depend on the readout axis 2 (mask shape (S, N0, N1, 1))")
            m = m[..., 0]
        m = m.astype(bool)
        self.nshots, self.n0, self.n1 = m.shape
        k0s, k1s, cnts, cols = [], [], [], []
        for s in range(self.nshots):
            k1, k0 = np.nonzero(m[s].T)  # row-major over (k1, k0): sorted by k1, then k0
            k0s.append(k0)
            k1s.append(k1)
            cnts.append(k0.size)
            cols.append(np.searchsorted(k1, np.arange(self.n1 + 1), side="left"))
        cnt = np.asarray(cnts, dtype=np.int64)
        off = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64)
        self.total = int(cnt.sum())
        if self.total >= 2**31:
            raise ValueError("too many samples")
        self.cnt_h = cnt
        self.off_h = off
        self.max_cnt = int(cnt.max()) if cnt.size else 0
        cat = lambda a: np.concatenate(a) if a else np.zeros(0, dtype=np.int64)
        k0a, k1a = cat(k0s), cat(k1s)
        if k0a.size == 0:
            k0a = k1a = np.zeros(1, dtype=np.int64)
        self.k0_h, self.k1_h = k0a[: self.total], k1a[: self.total]
        as_i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=device)
        self.sk0 = as_i32(k0a)
        self.sk1 = as_i32(k1a)
        self.off = as_i32(off)
        self.cnt = as_i32(cnt)
        self.colptr = as_i32(np.stack(cols) if cols else np.zeros((1, self.n1 + 1)))


class SenseProblem:
    """Encoding operator of one reconstruction level with its measured data.

    Inputs follow the reference's conventions: ksp and mps (C, N0, N1, N2)
    complex, shot_mask (S, N0, N1, 1) bool.  ``centered=True`` applies the
    reference's ifftshift on input (joint_recon.py:63-65); ``normalize=True``
    divides the data by max |ifftn(ksp)| (joint_recon.py:67-68).  P and M
    default to 1/(sum|S|^2 + 1e-3) and sum|S|^2 > 0.1 (joint_recon.py:83-85)."""

    def __init__(self, ksp, mps, shot_mask, kgrid, rkgrid, P=None, M=None, centered=False, normalize=False,
                 mem_budget=None):
        dev = _lib.require_gpu()
        self.device = dev
        self.geom = Geometry.from_grids(kgrid, rkgrid, dev)
        self.shape = self.geom.shape
        n0, n1, n2 = self.shape
        self.vol = n0 * n1 * n2
        ksp_d = as_c64(ksp, dev)
        mps_d = as_c64(mps, dev)
        if tuple(mps_d.shape[1:]) != self.shape or tuple(ksp_d.shape) != tuple(mps_d.shape):
            raise ValueError(f"ksp {tuple(ksp_d.shape)} / mps {tuple(mps_d.shape)} do not match the grid {self.shape}")
        self.ncoils = mps_d.shape[0]
        sh = "i" if centered else None
        # maps in plane-major layout
        self.maps = to_plane(mps_d, self.shape, sh, ncoils=self.ncoils)
        del mps_d
        mask = _to_numpy(shot_mask)
        if centered:  # (S, N0, N1[, 1]): shift the image axes only
            mask = np.fft.ifftshift(mask, axes=(1, 2))
        self.plan = ShotPlan(mask, dev)
        self.nshots = self.plan.nshots
        if (self.plan.n0, self.plan.n1) != (n0, n1):
            raise ValueError("shot mask shape does not match the image")
        # uncentred C-order k-space, readout inverse FFT -> hybrid space
        vol = self.vol
        kh = torch.empty_like(ksp_d)
        dims = (self.ncoils, *self.shape)
        permute(ksp_d, std_strides(0, self.shape), dims, kh, std_strides(0, self.shape), dims, shift_offsets(self.shape, sh))
        del ksp_d
        st = _lib.ll3((n1 * n2, n2, 1))
        nn = _lib.int3(self.shape)
        s = _lib.stream(dev)
        _lib.call("me_fft_axis", _lib.ptr(kh), vol, _lib.ptr(kh), vol, self.ncoils, nn, st, 2, 1, s)
        self.ymeas = torch.empty(max(1, self.plan.total * n2 * self.ncoils), dtype=torch.complex64, device=dev)
        _lib.call("me_gather_hybrid", _lib.ptr(kh), self.ncoils, nn, _lib.ptr(self.plan.sk0), _lib.ptr(self.plan.sk1),
                  _lib.ptr(self.plan.off), _lib.ptr(self.plan.cnt), self.nshots, self.plan.total, 1.0,
                  _lib.ptr(self.ymeas), s)
        self.max_norm = 1.0
        if normalize:  # max |ifftn(ksp)| over the full (unmasked) k-space
            for ax in (0, 1):
                _lib.call("me_fft_axis", _lib.ptr(kh), vol, _lib.ptr(kh), vol, self.ncoils, nn, st, ax, 1, s)
            bits = torch.zeros(1, dtype=torch.int32, device=dev)
            _lib.call("me_max_abs", _lib.ptr(kh), kh.numel(), _lib.ptr(bits), s)
            self.max_norm = float(bits.view(torch.float32).item())
            if self.max_norm > 0:
                self.scale_(self.ymeas, 1.0 / self.max_norm)
        del kh
        # preconditioner and support mask (plane-major)
        if P is None or M is None:
            Pp = torch.empty(vol, dtype=torch.float32, device=dev)
            Mp = torch.empty(vol, dtype=torch.uint8, device=dev)
            _lib.call("me_coil_ssq", _lib.ptr(self.maps), self.ncoils, vol, 1e-3, 0.1, _lib.ptr(Pp), _lib.ptr(Mp), s)
        self.P = real_to_plane(P, self.shape, torch.float32, dev) if P is not None else Pp
        if M is not None:
            Mt = real_to_plane(M, self.shape, torch.float32, dev)
            # the kernels apply M as a 0/1 support mask (the reference multiplies
            # by M, image_estimation.py:103, 129; its M is boolean, joint_recon.py:85)
            if not bool(((Mt == 0) | (Mt == 1)).all()):
                raise ValueError("M must be a binary (0/1 or boolean) support mask")
            self.M = Mt.to(torch.uint8)
        else:
            self.M = Mp
        # batching over shots
        budget = mem_budget or DEFAULT_MEM_BUDGET
        ncg = int(_lib.size("me_encode_coil_groups", self.ncoils))
        self.batch = int(max(1, min(self.nshots, budget // max(1, vol * 8 * ncg))))
        self.theta = torch.zeros((self.nshots, 6), dtype=torch.float64, device=dev)
        self.damp = torch.ones(self.nshots, dtype=torch.float64, device=dev)
        self.flags = torch.zeros(self.nshots, dtype=torch.uint8, device=dev)
        self._U = None
        self._W = None
        self._Y = None
        self._rhs_cache = None
        self.ncg = int(_lib.size("me_encode_coil_groups", self.ncoils))  # adjoint partial images per shot
        self.parts = min(self.ncg, 2)  # after folding, summed by the first adjoint pass
        self._vec_partials = torch.empty(_lib.size("me_vec_partials_size"), dtype=torch.float64, device=dev)
        self._cg_state = torch.zeros(10, dtype=torch.float64, device=dev)  # rule 0 (estimate_image_cg)
        nps = _lib.size("me_encode_partials_size", n2, self.ncoils, self.plan.max_cnt, n1)
        self._enc_partials = torch.empty(max(1, nps * self.batch), dtype=torch.float64, device=dev)

    # ------------------------------------------------------------ buffers
    def scale_(self, t, scale):
        n = t.numel()
        _lib.call("me_permute_shift", _lib.ptr(t), _ll4((0, 0, 0, 1)), _i4((1, 1, 1, n)), _lib.ptr(t),
                  _ll4((0, 0, 0, 1)), _i4((1, 1, 1, n)), _i4((0, 0, 0, 0)), float(scale), _lib.stream(t.device))

    def _ubuf(self):
        """Batch buffer: B transformed images, or B x ngroups adjoint partials."""
        if self._U is None:
            self._U = torch.empty((self.batch * self.ncg, self.vol), dtype=torch.complex64, device=self.device)
        return self._U

    def _wbuf(self):
        """Second batch buffer (B images): work volumes of the transforms."""
        if self._W is None:
            self._W = torch.empty((self.batch, self.vol), dtype=torch.complex64, device=self.device)
        return self._W

    def _ybuf(self):
        if self._Y is None:
            self._Y = torch.empty_like(self.ymeas)
        return self._Y

    def batches(self):
        for s0 in range(0, self.nshots, self.batch):
            yield s0, min(self.nshots, s0 + self.batch)

    def zeros_image(self):
        return torch.zeros(self.vol, dtype=torch.complex64, device=self.device)

    # ------------------------------------------------------------ kernels
    def _coef_ws(self, B):
        """Coefficient table of a transform call (B rows of 16 doubles); one
        persistent buffer, so the outer iteration allocates nothing and can
        be captured into a graph."""
        if getattr(self, "_coef", None) is None or self._coef.shape[0] < B:
            self._coef = torch.empty((max(B, self.batch), 16), dtype=torch.float64, device=self.device)
        return self._coef[:B]

    def _rigid(self, x, x_bs, out, theta, mode, flags=0, work=None, mask=None, x_parts=1):
        g = self.geom
        B = theta.shape[0]
        coef = self._coef_ws(B)
        _lib.call("me_rigid_transform", _lib.ptr(x), x_bs, int(x_parts), _lib.ptr(out), self.vol, _lib.ptr(work),
                  self.vol, int(B),
                  g.n_arg(), g.st_arg(True), *g.r_args(), _lib.ptr(theta), int(theta.stride(0)), _lib.ptr(coef),
                  int(mode), int(flags), _lib.ptr(mask), _lib.stream(self.device))

    def _deriv(self, x, out, theta, p_idx, work):
        g = self.geom
        B = theta.shape[0]
        coef = self._coef_ws(B)
        _lib.call("me_rigid_deriv", _lib.ptr(x), 0, _lib.ptr(out), self.vol, _lib.ptr(work), self.vol, int(B),
                  g.n_arg(), g.st_arg(True), *g.r_args(), _lib.ptr(theta), int(theta.stride(0)), _lib.ptr(coef),
                  int(p_idx), _lib.stream(self.device))

    def _deriv_chain(self, x, out, theta, p_idx, build):
        """Column p_idx of the LM Jacobian images from the shared rotation
        chain c1 = Rz x, c2 = Ry c1, c3 = Tr Rx c2 = T_d x (built when
        ``build``).  c3 lives in the work buffer, so columns 0-2 (one -i k_p
        pass over c3 each) come before 3-5."""
        g = self.geom
        B = theta.shape[0]
        if getattr(self, "_chain", None) is None:
            self._chain = torch.empty((2, self.batch, self.vol), dtype=torch.complex64, device=self.device)
        W = self._wbuf()
        coef = self._coef_ws(B)
        _lib.call("me_rigid_deriv_chain", _lib.ptr(x), 0, _lib.ptr(self._chain[0]), _lib.ptr(self._chain[1]),
                  _lib.ptr(W), self.vol, _lib.ptr(out), self.vol, _lib.ptr(W), self.vol, int(B), g.n_arg(),
                  g.st_arg(True), *g.r_args(), _lib.ptr(theta), int(theta.stride(0)), _lib.ptr(coef), int(p_idx),
                  int(bool(build)), _lib.stream(self.device))

    def _encode(self, u, u_bs, s0, s1, y=None, meas=False, norms=None, store_raw=False):
        """Gathered encodings of shots s0..s1 into y (nullable); with ``meas``
        the residual against the measured samples (or, with ``store_raw``, the
        encodings themselves while ``norms`` still gets the residual norms)."""
        pl = self.plan
        maxc = int(pl.cnt_h[s0:s1].max()) if s1 > s0 else 0
        _lib.call("me_encode_forward", _lib.ptr(u), u_bs, _lib.ptr(self.maps), self.ncoils, s1 - s0,
                  _lib.int3(self.shape), _lib.ptr(pl.sk0), _lib.ptr(pl.sk1), _lib.ptr(pl.off[s0:s1]),
                  _lib.ptr(pl.cnt[s0:s1]), maxc, _lib.ptr(y), _lib.ptr(self.ymeas) if meas else None,
                  _lib.ptr(self._enc_partials) if norms is not None else None, _lib.ptr(norms),
                  1 if store_raw else 0, _lib.stream(self.device))

    def _decode(self, y, s0, s1, out):
        pl = self.plan
        _lib.call("me_encode_adjoint", _lib.ptr(y), _lib.ptr(self.maps), self.ncoils, s1 - s0, _lib.int3(self.shape),
                  _lib.ptr(pl.sk0), _lib.ptr(pl.off[s0:s1]), _lib.ptr(pl.cnt[s0:s1]), _lib.ptr(pl.colptr[s0:s1]),
                  _lib.ptr(out), self.vol, _lib.stream(self.device))

    # ------------------------------------------------------------ operators
    def forward(self, x, theta=None, meas=False, y=None):
        """Gathered E_s x (or residual E_s x - y_s) for every shot."""
        theta = self.theta if theta is None else theta
        U = self._ubuf()
        y = self._ybuf() if y is None else y
        for s0, s1 in self.batches():
            self._rigid(x, 0, U, theta[s0:s1], 0)
            self._encode(U, self.vol, s0, s1, y=y, meas=meas)
        return y

    def _adjoint_sum(self, parts, s0, s1, out, theta):
        """out (+)= M * sum_{s in batch} T_s^H (sum of the shot's partial images)."""
        flags = 1 | 2 | (4 if s0 > 0 else 0)
        self._rigid(parts, self.vol * self.ncg, out, theta[s0:s1], 1, flags=flags, work=self._wbuf(), mask=self.M,
                    x_parts=self.parts)

    def normal(self, p, out, theta=None, y=None, norms=None, residual=False):
        """out = M * sum_s T_s^H E_s^H E_s T_s p  (image_estimation.py:106-129).

        The gathered encodings E_s T_s p are left in ``y`` (default: the
        problem's sample buffer); with ``norms`` (nshots FP64) the per-shot
        ||E_s T_s p - y_s||^2 come out of the same encoding.  ``residual``:
        the encodings are taken against the measured data, so ``y`` holds
        E_s T_s p - y_s and out = M * sum_s T_s^H E_s^H (E_s T_s p - y_s), which
        is A p - b with b the right-hand side of image_estimation.py:92-103
        (the encoding and the sums are linear), without a separate b."""
        theta = self.theta if theta is None else theta
        U = self._ubuf()
        Y = self._ybuf() if y is None else y
        for s0, s1 in self.batches():
            self._rigid(p, 0, U, theta[s0:s1], 0)
            if residual:
                self._encode(U, self.vol, s0, s1, y=Y, meas=True,
                             norms=None if norms is None else norms[s0:s1])
            elif norms is None:
                self._encode(U, self.vol, s0, s1, y=Y)
            else:
                self._encode(U, self.vol, s0, s1, y=Y, meas=True, norms=norms[s0:s1], store_raw=True)
            self._decode(Y, s0, s1, U)
            self._adjoint_sum(U, s0, s1, out, theta)
        return out

    def rhs(self, theta=None, out=None, cache=True):
        """b = M * sum_s T_s^H E_s^H y_s  (image_estimation.py:92-103).
        E_s^H y_s does not depend on the motion and is cached when it fits."""
        theta = self.theta if theta is None else theta
        out = self.zeros_image() if out is None else out
        if cache and self._rhs_cache is None and self.nshots * self.ncg * self.vol * 8 <= int(16 * 2**30):
            self._rhs_cache = torch.empty((self.nshots * self.ncg, self.vol), dtype=torch.complex64,
                                          device=self.device)
            for s0, s1 in self.batches():
                self._decode(self.ymeas, s0, s1, self._rhs_cache[s0 * self.ncg:s1 * self.ncg])
        U = self._ubuf()
        for s0, s1 in self.batches():
            if cache and self._rhs_cache is not None:
                src = self._rhs_cache[s0 * self.ncg:s1 * self.ncg]
            else:
                self._decode(self.ymeas, s0, s1, U)
                src = U
            self._adjoint_sum(src, s0, s1, out, theta)
        return out

    def cg(self, x, b, max_iter, tol=1e-6, check_every=1, theta=None, track=None, norms0=None):
        """PCG of image_estimation.py:130-165 (x updated in place).

        ``b=None`` takes the right-hand side of image_estimation.py:92-103
        implicitly: the initial residual r0 = b - A x0 is formed as
        -M sum_s T_s^H E_s^H (E_s T_s x0 - y_s) in one normal operator, which
        saves b's own adjoint transforms.

        ``track`` (a sample-sized buffer) follows E T x through the steps: the
        initial normal operator writes E T x0 there (E T x0 - y with
        ``b=None``) and every step adds alpha E T p, the encodings the normal
        operator of p already made, so it ends as E T x (or E T x - y) of the
        final x.  ``norms0`` (nshots FP64) receives ||E_s T_s x0 - y_s||^2 of
        the initial x from the same first encoding."""
        n = self.vol
        if getattr(self, "_cgv", None) is None:
            self._cgv = torch.empty((3, n), dtype=torch.complex64, device=self.device)
        Ap, r, p = self._cgv[0], self._cgv[1], self._cgv[2]
        st = self._cg_state
        s = _lib.stream(self.device)
        residual = b is None
        self.normal(x, Ap, theta, y=track, norms=norms0, residual=residual)
        if residual:  # r = 0 - (A x0 - b)
            if getattr(self, "_zero_img", None) is None:
                self._zero_img = self.zeros_image()
            b = self._zero_img
        self._track_is_residual = residual and track is not None
        _lib.call("me_cg_init", _lib.ptr(b), _lib.ptr(Ap), _lib.ptr(r), _lib.ptr(p), _lib.ptr(self.P), n,
                  _lib.ptr(st), _lib.ptr(self._vec_partials), s)
        Y = self._ybuf()
        ny = Y.numel() if track is not None else 0
        for it in range(max_iter):
            self.normal(p, Ap, theta, y=Y)
            _lib.call("me_cg_step", _lib.ptr(x), _lib.ptr(r), _lib.ptr(p), _lib.ptr(Ap), _lib.ptr(self.P), n,
                      float(tol), _lib.ptr(st), _lib.ptr(self._vec_partials), _lib.ptr(track),
                      _lib.ptr(Y) if track is not None else None, ny, s)
            if check_every and (it + 1) % check_every == 0 and it + 1 < max_iter:
                if float(st[CG_DONE].item()) != 0.0:
                    break
        return x

    def track_buffer(self):
        """Sample-sized buffer for cg(track=...) / lm_step(tracked=...)."""
        if getattr(self, "_R", None) is None:
            self._R = torch.empty_like(self.ymeas)
        return self._R

    def objective_per_shot(self, x, theta=None):
        theta = self.theta if theta is None else theta
        U = self._ubuf()
        norms = torch.empty(self.nshots, dtype=torch.float64, device=self.device)
        for s0, s1 in self.batches():
            self._rigid(x, 0, U, theta[s0:s1], 0)
            self._encode(U, self.vol, s0, s1, y=None, meas=True, norms=norms[s0:s1])
        return norms

    def objective(self, x, theta=None):
        """sum_s ||M_s (F S T_s x - y)||^2  (joint_recon.py:108-115)."""
        return float(self.objective_per_shot(x, theta).sum().item())

    def lm_step(self, x, tol=1e-6, record=None, tracked=False, use_ga=False, out=None):
        """One LM update of every shot (transform_estimation.py:39-150).
        Updates self.theta/damp/flags in place, returns T(mean) x (written to
        ``out`` when given, which may be x itself).

        ``tracked``: the track buffer already holds E T x (or, after
        cg(b=None, track=...), E T x - y) for this x and theta, so the residual
        is formed from it instead of a new transform and encoding.  ``use_ga``: the reference's experimental
        geodesic acceleration (transform_estimation.py:73-121): six more
        columns at theta + 1e-4 v per shot and a second 6x6 solve."""
        dev = self.device
        nY = self.ymeas.numel()
        R = self.track_buffer()
        if getattr(self, "_J", None) is None:
            self._J = torch.empty((6, nY), dtype=torch.complex64, device=dev)
            self._stats = torch.empty((self.nshots, 28), dtype=torch.float64, device=dev)
            self._e1 = torch.empty(self.nshots, dtype=torch.float64, device=dev)
            self._cand = torch.empty((self.nshots, 6), dtype=torch.float64, device=dev)
            self._delta = torch.empty((self.nshots, 6), dtype=torch.float64, device=dev)
            self._acc = torch.empty(self.nshots, dtype=torch.uint8, device=dev)
            self._lm_partials = torch.empty(_lib.size("me_lm_partials_size", self.batch), dtype=torch.float64,
                                            device=dev)
        U, W = self._ubuf(), self._wbuf()
        s = _lib.stream(dev)
        n0, n1, n2 = self.shape
        th = self.theta
        if tracked and not getattr(self, "_track_is_residual", False):
            _lib.call("me_csub", _lib.ptr(R), _lib.ptr(R), _lib.ptr(self.ymeas), nY, s)
        self._track_is_residual = False  # the buffer is reused below
        for s0, s1 in self.batches():
            B = s1 - s0
            if not tracked:
                self._rigid(x, 0, U, th[s0:s1], 0)
                self._encode(U, self.vol, s0, s1, y=R, meas=True)
            for p_idx in range(6):
                self._deriv_chain(x, U, th[s0:s1], p_idx, p_idx == 0)
                self._encode(U, self.vol, s0, s1, y=self._J[p_idx])
            _lib.call("me_lm_gram", _lib.ptr(R), _lib.ptr(self._J), nY, _lib.ptr(self.plan.off[s0:s1]),
                      _lib.ptr(self.plan.cnt[s0:s1]), n2 * self.ncoils, B, _lib.ptr(self._lm_partials),
                      _lib.ptr(self._stats[s0:s1]), s)
            _lib.call("me_lm_solve", _lib.ptr(self._stats[s0:s1]), _lib.ptr(self.damp[s0:s1]), _lib.ptr(th[s0:s1]), 6,
                      _lib.ptr(self._cand[s0:s1]), _lib.ptr(self._delta[s0:s1]), B, s)
            if use_ga:
                self._ga_candidates(x, s0, s1)
            self._rigid(x, 0, U, self._cand[s0:s1], 0)
            self._encode(U, self.vol, s0, s1, y=None, meas=True, norms=self._e1[s0:s1])
            _lib.call("me_lm_accept", _lib.ptr(self._stats[s0:s1]), _lib.ptr(self._e1[s0:s1]), _lib.ptr(th[s0:s1]), 6,
                      _lib.ptr(self._cand[s0:s1]), _lib.ptr(self.damp[s0:s1]), _lib.ptr(self.flags[s0:s1]),
                      _lib.ptr(self._acc[s0:s1]), float(tol), B, s)
        if record is not None:
            record.append(dict(e0=self._stats[:, 0].cpu().numpy().copy(), e1=self._e1.cpu().numpy().copy(),
                               accepted=self._acc.cpu().numpy().astype(bool), delta=self._delta.cpu().numpy().copy()))
        if getattr(self, "_mean", None) is None:
            self._mean = torch.empty((1, 6), dtype=torch.float64, device=dev)
        mean = self._mean
        _lib.call("me_lm_mean", _lib.ptr(th), 6, self.nshots, _lib.ptr(mean), s)
        if out is None:
            out = torch.empty(self.vol, dtype=torch.complex64, device=dev)
        self._rigid(x, 0, out, mean, 0)  # in place when out is x
        return out

    def _ga_candidates(self, x, s0, s1, h=GA_STEP):
        """Replace the candidates of shots s0..s1 by their geodesically
        accelerated ones (transform_estimation.py:73-121).

        J'v, the derivative of the Jacobian along v = -delta, is estimated by
        a central difference at h = 1e-3 of the derivative images,
        sum_p v_p (dT_p(theta + h v) x - dT_p(theta - h v) x) / (2h), which is
        encoded once (the encoding is linear).  The reference's forward
        difference at eps = 1e-4 of FP32 Jacobian columns
        (transform_estimation.py:76-98) estimates the same quantity, but its
        rounding noise is about ten times larger; the two differ from each
        other by the forward difference's O(eps) term."""
        dev = self.device
        nY = self.ymeas.numel()
        B = s1 - s0
        n2 = self.shape[2]
        if getattr(self, "_Jpv", None) is None:
            self._Jpv = torch.empty(nY, dtype=torch.complex64, device=dev)
            self._stats_dir = torch.empty((self.nshots, 28), dtype=torch.float64, device=dev)
            self._ga_img = torch.empty((self.batch, self.vol), dtype=torch.complex64, device=dev)
        s = _lib.stream(dev)
        U = self._ubuf()
        D = self._ga_img
        for sign in (1.0, -1.0):
            th_h = self.theta[s0:s1] - (sign * h) * self._delta[s0:s1]  # theta +- h v
            for p_idx in range(6):
                self._deriv_chain(x, U, th_h, p_idx, p_idx == 0)
                _lib.call("me_lm_ga_accumulate", _lib.ptr(U), self.vol, self.vol, B, _lib.ptr(self._delta[s0:s1]),
                          p_idx, sign / (2.0 * h), int(sign < 0 or p_idx > 0), _lib.ptr(D), self.vol, s)
        pl = self.plan
        self._encode(D, self.vol, s0, s1, y=self._Jpv)
        _lib.call("me_lm_gram", _lib.ptr(self._Jpv), _lib.ptr(self._J), nY, _lib.ptr(pl.off[s0:s1]),
                  _lib.ptr(pl.cnt[s0:s1]), n2 * self.ncoils, B, _lib.ptr(self._lm_partials),
                  _lib.ptr(self._stats_dir[s0:s1]), s)
        _lib.call("me_lm_ga_solve", _lib.ptr(self._stats[s0:s1]), _lib.ptr(self._stats_dir[s0:s1]),
                  _lib.ptr(self.damp[s0:s1]), _lib.ptr(self.theta[s0:s1]), 6, _lib.ptr(self._delta[s0:s1]),
                  _lib.ptr(self._cand[s0:s1]), B, s)

    def mask_(self, x):
        _lib.call("me_mask_mul", _lib.ptr(x), _lib.ptr(self.M), x.numel(), _lib.stream(self.device))
        return x

    # ------------------------------------------------------------ layout helpers
    def image_to_plane(self, x, shift=None):
        return to_plane(as_c64(x, self.device), self.shape, shift)

    def image_from_plane(self, x, shift=None, scale=1.0):
        return from_plane(x, self.shape, shift, scale)
