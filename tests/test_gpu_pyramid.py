"""GPU parity of the pyramid driver (pyramid.py:24-118) against the FP64
oracle at small sizes: one coarse level (x2 along axes 0 and 1, DISORDER)
with JointRecon, then CG-only reconstruction at full resolution; and the
reference's default three levels (x4, x2, then CG).  Relative L2 1e-4 on the
image, 1e-3 (degrees, voxels) on motion."""
import numpy as np
import pytest

from oracle import motion_oracle as O

pytestmark = pytest.mark.gpu
SHAPE = (32, 32, 8)
NC = 4


@pytest.mark.parametrize("tune", [False, True])
def test_pyramid_matches_oracle(me, tune):
    """With ``tune`` the autotuners pick each level's kernel variants first;
    the result must not depend on it."""
    sim = me.simulation
    gt = sim.phantom(SHAPE, seed=0).numpy().astype(np.complex128)
    mps = sim.coil_maps(SHAPE, NC, seed=1).numpy().astype(np.complex128)
    A, _ = sim.disorder_sampling(SHAPE, (4, 4), 2)
    t_true = sim.motion_uniform(A.shape[0], rot_deg=2.0, trans_vox=1.0, seed=2)
    ksp = O.generate_corrupted_kspace(gt, mps, A, t_true)
    kw = dict(n_spatial_levels=2, n_joint_iters=1, final_cg_iter=20)
    x_got, t_got = me.pyramid_reconstruction(ksp, mps, A, A.shape[0], compute_objective=False, tune=tune, **kw)
    x_ref, t_ref = O.pyramid_reconstruction(ksp, mps, A, A.shape[0], **kw)
    d = np.abs(np.asarray(t_got) - t_ref)
    assert max(d[:, :3].max(), np.degrees(d[:, 3:]).max()) < 1e-3
    assert np.linalg.norm(x_got - x_ref) / np.linalg.norm(x_ref) < 1e-4
    assert np.abs(t_ref).max() > 1e-3  # the coarse level did estimate motion
    me._lib.load().me_line_tuning_clear()
    me._lib.load().me_encode_tuning_clear()


def test_three_level_pyramid_matches_oracle(me):
    """The reference's default depth (n_spatial_levels=3, pyramid.py:24):
    JointRecon at x4 and x2 (lines of 16 and 32), the motion carried
    unscaled between them, then CG at full resolution."""
    shape = (64, 64, 8)
    sim = me.simulation
    gt = sim.phantom(shape, seed=3).numpy().astype(np.complex128)
    mps = sim.coil_maps(shape, NC, seed=4).numpy().astype(np.complex128)
    A, _ = sim.disorder_sampling(shape, (4, 4), 2)
    t_true = sim.motion_uniform(A.shape[0], rot_deg=2.0, trans_vox=1.0, seed=5)
    ksp = O.generate_corrupted_kspace(gt, mps, A, t_true)
    kw = dict(n_spatial_levels=3, n_joint_iters=1, final_cg_iter=10)
    x_got, t_got = me.pyramid_reconstruction(ksp, mps, A, A.shape[0], compute_objective=False, **kw)
    x_ref, t_ref = O.pyramid_reconstruction(ksp, mps, A, A.shape[0], **kw)
    d = np.abs(np.asarray(t_got) - t_ref)
    assert max(d[:, :3].max(), np.degrees(d[:, 3:]).max()) < 1e-3
    assert np.linalg.norm(x_got - x_ref) / np.linalg.norm(x_ref) < 1e-4
    assert np.abs(t_ref).max() > 1e-3
