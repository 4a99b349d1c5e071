"""Data-consistency loss with a hand-written backward.

``data_consistency_loss(x, theta, problem)`` returns

    L(x, theta) = 0.5 * sum_s || M_s F (S * T(theta_s) x) - y_s ||^2

(the per-shot residual of transform_estimation.py:41-43 summed over shots, T
the ``apply`` of transform.py:263-318, i.e. Tr Rz Ry Rx) as a differentiable
torch scalar.  The backward is written out, not traced:

* image:  dL/dRe(x) + i dL/dIm(x) = sum_s T_s^H E_s^H r_s   (adjoint encoding of
  the gathered residual, then the batched adjoint transform summed over shots);
* motion: dL/dtheta_{s,p} = Re < E_s^H r_s, dT_s/dtheta_p x >, with the six
  Jacobian columns of the apply order built from shared rotation chains
  (c1 = Rx x, c2 = Tr_z Ry c1, u = T x), each reduced per shot in FP64 by
  its own last line pass instead of being stored (me_rigid_deriv_apply_dot).

The sheared intermediates (u, c1, c2) are recomputed in the backward by
default (u through the chain the columns start from, so the chain is built
once); ``recompute=False`` keeps u, c1 and c2 from the forward instead (three
volumes per shot), whose transform then runs through the same chain.  Inputs follow ``estimate_transform``'s frame (uncentred
ksp/mps/masks) through a :class:`SenseProblem`.
"""
from __future__ import annotations

import torch

from . import _lib
from .encoding import SenseProblem
from .utils import as_c64


def _bufs(pb, n):
    """n batch-sized volume buffers cached on the problem."""
    b = getattr(pb, "_grad_bufs", None)
    if b is None or b.shape[0] < n:
        b = torch.empty((n, pb.batch, pb.vol), dtype=torch.complex64, device=pb.device)
        pb._grad_bufs = b
    return b


def _deriv_apply(pb, x, u, u_bs, c1, c2, out, work, theta, p_idx, build):
    g = pb.geom
    B = theta.shape[0]
    coef = torch.empty((B, 16), dtype=torch.float64, device=pb.device)
    _lib.call("me_rigid_deriv_apply", _lib.ptr(x), 0, _lib.ptr(u), u_bs, _lib.ptr(c1), _lib.ptr(c2), pb.vol,
              _lib.ptr(out), pb.vol, _lib.ptr(work), pb.vol, int(B), g.n_arg(), g.st_arg(True), *g.r_args(),
              _lib.ptr(theta), int(theta.stride(0)), _lib.ptr(coef), int(p_idx), int(bool(build)),
              _lib.stream(pb.device))


def _apply_chain(pb, x, c1, c2, u, theta):
    """u = T x (bit for bit me_rigid_transform's apply) keeping the chain
    c1 = Rx x, c2 = Tr_z Ry c1 that the Jacobian columns start from."""
    g = pb.geom
    B = theta.shape[0]
    coef = torch.empty((B, 16), dtype=torch.float64, device=pb.device)
    _lib.call("me_rigid_apply_chain", _lib.ptr(x), 0, _lib.ptr(c1), _lib.ptr(c2), pb.vol, _lib.ptr(u), pb.vol,
              int(B), g.n_arg(), g.st_arg(True), *g.r_args(), _lib.ptr(theta), int(theta.stride(0)),
              _lib.ptr(coef), _lib.stream(pb.device))


def _deriv_apply_dot(pb, x, u, u_bs, c1, c2, w, y, work, theta, p_idx, build, out, out_stride):
    """out[b * out_stride] = Re <w_b, dT_b/dtheta_p x> (FP64): the column's
    last line pass reduces against w instead of storing the column."""
    g = pb.geom
    B = theta.shape[0]
    coef = torch.empty((B, 16), dtype=torch.float64, device=pb.device)
    n = int(_lib.size("me_rigid_dot_partials_size", int(B), g.n_arg(), g.st_arg(True)))
    part = getattr(pb, "_rdot_partials", None)
    if part is None or part.numel() < n:
        part = torch.empty(max(n, 1), dtype=torch.float64, device=pb.device)
        pb._rdot_partials = part
    _lib.call("me_rigid_deriv_apply_dot", _lib.ptr(x), 0, _lib.ptr(u), u_bs, _lib.ptr(c1), _lib.ptr(c2), pb.vol,
              _lib.ptr(w), pb.vol, _lib.ptr(y), pb.vol, _lib.ptr(work), pb.vol, int(B), g.n_arg(), g.st_arg(True),
              *g.r_args(), _lib.ptr(theta), int(theta.stride(0)), _lib.ptr(coef), int(p_idx), int(bool(build)),
              _lib.ptr(part), _lib.ptr(out), int(out_stride), _lib.stream(pb.device))


def _dots(pb, a, c, B, out, out_stride):
    """out[b * out_stride] = Re <a_b, c_b> (FP64) for b < B."""
    n = int(_lib.size("me_dot_batched_partials_size", B))
    part = getattr(pb, "_dot_partials", None)
    if part is None or part.numel() < n:
        part = torch.empty(max(n, 1), dtype=torch.float64, device=pb.device)
        pb._dot_partials = part
    _lib.call("me_dot_re_batched", _lib.ptr(a), pb.vol, _lib.ptr(c), pb.vol, pb.vol, int(B), _lib.ptr(part),
              _lib.ptr(out), int(out_stride), _lib.stream(pb.device))


class DataConsistencyLoss(torch.autograd.Function):
    """torch.autograd.Function for ``data_consistency_loss``."""

    @staticmethod
    def forward(ctx, x, theta, problem, recompute=True):
        pb = problem
        dev = pb.device
        xp = pb.image_to_plane(x.detach())
        th = theta.detach().to(device=dev, dtype=torch.float64).reshape(pb.nshots, 6).contiguous()
        R = torch.empty_like(pb.ymeas)
        norms = torch.empty(pb.nshots, dtype=torch.float64, device=dev)
        # recompute=False keeps u = T x and its chain (c1, c2) for the backward
        K = None if recompute else torch.empty((3, pb.nshots, pb.vol), dtype=torch.complex64, device=dev)
        ub = _bufs(pb, 1)[0]
        for s0, s1 in pb.batches():
            if recompute:
                dst = ub
                pb._rigid(xp, 0, dst, th[s0:s1], 0)
            else:
                dst = K[0, s0:s1]
                _apply_chain(pb, xp, K[1, s0:s1], K[2, s0:s1], dst, th[s0:s1])
            pb._encode(dst, pb.vol, s0, s1, y=R, meas=True, norms=norms[s0:s1])
        ctx.problem = pb
        ctx.recompute = recompute
        ctx.x_dtype = x.dtype
        ctx.theta_dtype = theta.dtype
        ctx.theta_shape = theta.shape
        if recompute:
            ctx.save_for_backward(xp, th, R)
        else:
            ctx.save_for_backward(xp, th, R, K)
        return 0.5 * norms.sum()

    @staticmethod
    def backward(ctx, grad_out):
        pb = ctx.problem
        saved = ctx.saved_tensors
        xp, th, R = saved[:3]
        Kst = None if ctx.recompute else saved[3]
        need_x, need_t = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        gx = pb.zeros_image()
        gt = torch.zeros((pb.nshots, 6), dtype=torch.float64, device=pb.device)
        bufs = _bufs(pb, 5)
        W, U, c1, c2, col = bufs[0], bufs[1], bufs[2], bufs[3], bufs[4]
        work = pb._wbuf()
        for s0, s1 in pb.batches():
            B = s1 - s0
            pb._decode(R, s0, s1, W)  # E_s^H r_s, one image per shot
            if need_x:
                pb._rigid(W, pb.vol * pb.ncg, gx, th[s0:s1], 1, flags=1 | (4 if s0 > 0 else 0), work=work,
                          x_parts=pb.parts)
            if need_t:
                if Kst is None:  # u = T x through the chain the columns start from
                    u, k1, k2 = U, c1, c2
                    _apply_chain(pb, xp, k1, k2, u, th[s0:s1])
                else:
                    u, k1, k2 = Kst[0, s0:s1], Kst[1, s0:s1], Kst[2, s0:s1]
                for p in range(6):
                    _deriv_apply_dot(pb, xp, u, pb.vol, k1, k2, W, col, work, th[s0:s1], p, False,
                                     gt[s0:s1, p:], 6)
        g = grad_out.to(torch.float64)
        grad_x = grad_t = None
        if need_x:
            grad_x = pb.image_from_plane(gx).mul_(g.to(torch.float32)).to(ctx.x_dtype)
        if need_t:
            grad_t = (gt * g).to(ctx.theta_dtype).reshape(ctx.theta_shape)
        return grad_x, grad_t, None, None


def data_consistency_loss(x, theta, problem, recompute=True):
    """0.5 sum_s ||M_s F(S T(theta_s) x) - y_s||^2 with a hand-written backward.

    x: (N0, N1, N2) complex image (torch, requires_grad as needed); theta:
    (nshots, 6) real [q1, q2, q3, t1, t2, t3] per shot; problem: a
    :class:`SenseProblem` holding the maps, masks and measured data.  Returns a
    float64 scalar tensor."""
    if not isinstance(problem, SenseProblem):
        raise TypeError("problem must be a SenseProblem")
    if x.dtype != torch.complex64:
        x = x.to(torch.complex64) if isinstance(x, torch.Tensor) else as_c64(x, problem.device)
    return DataConsistencyLoss.apply(x, theta, problem, recompute)
