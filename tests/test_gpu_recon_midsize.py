"""JointRecon against the FP64 oracle at a size whose rows use the same
16-thread line kernels as the 256 x 256 x 192 workload (N1 = 128) and whose
columns are not a power of two (N0 = 48): 4 coils, 8 DISORDER shots of 384
samples per readout plane, two outer iterations of 3 CG steps and one LM
step.  Relative L2 1e-4 on the image and the objective history, 1e-3
(degrees, voxels) on motion, damping identical."""
import numpy as np
import pytest

from oracle import motion_oracle as O

pytestmark = pytest.mark.gpu
SHAPE = (48, 128, 16)
NC = 4


def test_joint_recon_midsize_matches_oracle(me):
    sim = me.simulation
    gt = sim.phantom(SHAPE, seed=0).numpy().astype(np.complex128)
    mps = sim.coil_maps(SHAPE, NC, seed=1).numpy().astype(np.complex128)
    A, _ = sim.disorder_sampling(SHAPE, (4, 4), 2)
    t_true = sim.motion_uniform(A.shape[0], rot_deg=3.0, trans_vox=1.5, seed=2)
    ksp = O.generate_corrupted_kspace(gt, mps, A, t_true)
    kg, rkg = me.compute_transform_grids_voxel(SHAPE, (1, 1, 1))
    okg, orkg = O.compute_transform_grids_voxel(SHAPE, (1, 1, 1))
    kw = dict(max_cg_iter=3, max_joint_iter=2)
    app = me.JointRecon(ksp, mps, A, kg, rkg, **kw)
    x_got, t_got = app.run()
    ref = O.JointRecon(ksp, mps, A, okg, orkg, **kw)
    x_ref, t_ref = ref.run()
    x_got = x_got.cpu().numpy()
    assert np.linalg.norm(x_got - x_ref) / np.linalg.norm(x_ref) < 1e-4
    d = np.abs(t_got.cpu().numpy() - t_ref)
    assert max(d[:, :3].max(), np.degrees(d[:, 3:]).max()) < 1e-3
    assert np.allclose(app.objective_history, ref.objective_history, rtol=1e-4, atol=0)
    assert np.array_equal(app.damp.cpu().numpy(), ref.damp)
