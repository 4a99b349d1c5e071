"""Edge cases of the joint reconstruction against the FP64 oracle
(joint_recon.py:88-115, transform_estimation.py:39-144).

* Shots without samples (a caller-supplied mask whose rows for shots 0 and 3
  are all false): their encodings are empty, their LM residual, gradient and
  Gram are zero, so delta = 0, e1 = e0 = 0 and the update is rejected
  (damping x 1.5) while the other shots proceed; the mean re-registration
  still moves every shot.
* One shot holding the whole sampling pattern.

Both at 32 x 24 x 16 with 4 coils, two outer iterations of 3 CG + LM with the
objective: image relative L2 1e-4, motion 1e-3 (degrees, voxels), objective
relative 1e-4, damping identical."""
import numpy as np
import pytest

from oracle import motion_oracle as O

pytestmark = pytest.mark.gpu
SHAPE = (32, 24, 16)
NC = 4


def _problem(me, mask_fn):
    sim = me.simulation
    gt = sim.phantom(SHAPE, seed=0).numpy().astype(np.complex128)
    mps = sim.coil_maps(SHAPE, NC, seed=1).numpy().astype(np.complex128)
    A, _ = sim.disorder_sampling(SHAPE, (4, 4), 2)
    A = mask_fn(A)
    t_true = sim.motion_uniform(A.shape[0], rot_deg=3.0, trans_vox=1.5, seed=2)
    ksp = O.generate_corrupted_kspace(gt, mps, A, t_true)
    rng = np.random.default_rng(3)
    ksp = ksp + 1e-3 * np.abs(ksp).max() * (rng.standard_normal(ksp.shape) + 1j * rng.standard_normal(ksp.shape))
    return mps, A, ksp


def _compare(me, mps, A, ksp):
    kg, rkg = me.compute_transform_grids_voxel(SHAPE, (1, 1, 1))
    okg, orkg = O.compute_transform_grids_voxel(SHAPE, (1, 1, 1))
    app = me.JointRecon(ksp, mps, A, kg, rkg, max_cg_iter=3, max_joint_iter=2)
    x, t = app.run()
    ref = O.JointRecon(ksp, mps, A, okg, orkg, max_cg_iter=3, max_joint_iter=2)
    xr, tr = ref.run()
    x, t = x.cpu().numpy(), t.cpu().numpy()
    d = np.abs(t - tr)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-4
    assert max(d[:, :3].max(), np.degrees(d[:, 3:]).max()) < 1e-3
    assert np.allclose(app.objective_history, ref.objective_history, rtol=1e-4, atol=0)
    damp = app.damp.cpu().numpy()
    assert np.array_equal(damp, ref.damp)
    return t, damp


def test_empty_shots_match_oracle(me):
    def drop(A):
        A = A.copy()
        A[[0, 3]] = False
        return A

    mps, A, ksp = _problem(me, drop)
    assert A.shape[0] == 8 and not A[0].any() and not A[3].any() and A[1].any()
    t, damp = _compare(me, mps, A, ksp)
    # the empty shots were rejected twice (1 * 1.5 * 1.5); the others moved
    assert damp[0] == damp[3] == 2.25
    assert np.abs(t[1]).max() > 1e-4


def test_single_shot_matches_oracle(me):
    mps, A, ksp = _problem(me, lambda A: A.any(axis=0, keepdims=True))
    assert A.shape[0] == 1
    _compare(me, mps, A, ksp)
