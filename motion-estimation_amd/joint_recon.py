"""Joint image/motion reconstruction driver (joint_recon.py:10-115) on the GPU."""
from __future__ import annotations

import numpy as np
import torch

from .encoding import SenseProblem
from .utils import as_f64


class JointRecon:
    """Alternating CG image update and LM motion update (joint_recon.py:10-115).

    Same constructor as the reference: ksp/mps (ncoils, N0, N1, N2) and
    shot_mask (nshots, N0, N1, 1) in the CENTRED frame (they are uncentred
    here, joint_recon.py:63-65), the data normalised by max |ifftn(ksp)|
    (joint_recon.py:67-68), x0 = 0, t0 or zeros, damping 1 per shot, P/M from
    the maps.  ``run()`` returns (fftshift(x * max_norm), t) and fills
    ``transform_history`` (numpy) and ``objective_history`` (numpy).
    ``compute_objective=False`` skips the per-iteration objective (one extra
    forward encoding, joint_recon.py:102) for timing runs.  ``t``, ``damp`` and
    ``convergence_flags`` are the device state; ``x`` is the image in the
    library's plane-major layout (``image()`` gives the reference layout).
    """

    def __init__(self, ksp, mps, shot_mask, kgrid, rkgrid, t0=None, max_cg_iter=3, max_nm_iter=1,
                 max_joint_iter=100, xp=np, compute_objective=True, cg_check_every=0):
        self.kgrid, self.rkgrid = kgrid, rkgrid
        self.max_cg_iter = max_cg_iter
        self.max_nm_iter = max_nm_iter
        self.max_joint_iter = max_joint_iter
        self.compute_objective = compute_objective
        self.cg_check_every = cg_check_every
        self.num_shots = len(shot_mask)
        self.img_shape = tuple(ksp.shape[1:])
        self.problem = SenseProblem(ksp, mps, shot_mask, kgrid, rkgrid, centered=True, normalize=True)
        self.max_norm = self.problem.max_norm
        self.x = self.problem.zeros_image()
        if t0 is not None:
            self.problem.theta.copy_(as_f64(t0, self.problem.device).reshape(self.num_shots, 6))

    @property
    def t(self):
        return self.problem.theta

    @property
    def damp(self):
        """LM damping per shot (joint_recon.py:81), FP64 device tensor."""
        return self.problem.damp

    @property
    def convergence_flags(self):
        """Per-shot flags of transform_estimation.py:139-140 (written, never
        read, as in the reference), uint8 device tensor."""
        return self.problem.flags

    def step(self, prev_objective=None):
        """One outer iteration (joint_recon.py:92-100).

        The CG forms its initial residual b - A x0 from one normal operator of
        E T x0 - y (no separate right-hand side) and tracks E T x - y, so the
        first LM residual needs no new transform and encoding.
        ``prev_objective`` (nshots FP64 device tensor, or None) receives the
        per-shot objective terms of the state this step starts from, i.e. the
        previous iteration's objective (joint_recon.py:102), computed from the
        CG's first encoding of that same state."""
        pb = self.problem
        track = pb.track_buffer() if self.max_nm_iter > 0 else None
        # b = None: r0 = b - A x0 from one normal operator of the residual
        # (estimate_image_cg's b, image_estimation.py:92-103, implicitly)
        pb.cg(self.x, None, self.max_cg_iter, 1e-6, check_every=self.cg_check_every, track=track,
              norms0=prev_objective)
        for i in range(self.max_nm_iter):
            pb.lm_step(self.x, tracked=(i == 0), out=self.x)
            pb.mask_(self.x)

    def objective(self):
        return self.problem.objective(self.x)

    def run(self):
        self.transform_history = []
        pending = []
        for it in range(self.max_joint_iter):
            buf = None
            if self.compute_objective and it > 0:
                buf = torch.empty(self.num_shots, dtype=torch.float64, device=self.problem.device)
                pending.append(buf)
            self.step(prev_objective=buf)
            self.transform_history.append(self.t.detach().cpu().numpy().copy())
        if self.compute_objective and self.max_joint_iter > 0:
            pending.append(self.problem.objective_per_shot(self.x))
        self.objective_history = [float(o.sum().item()) for o in pending]
        self.transform_history = np.stack(self.transform_history) if self.transform_history else np.zeros((0, self.num_shots, 6))
        self.objective_history = np.array(self.objective_history)
        return self.image(), self.t

    def image(self):
        """fftshift(x * max_norm) in the reference layout (joint_recon.py:106)."""
        return self.problem.image_from_plane(self.x, shift="f", scale=self.max_norm)
