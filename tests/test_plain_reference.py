"""The plain torch.fft restatement of the reference's path (bench.py:
plain_transform, plain_deriv, PlainJointRecon; dense masks, coil-batched 3-D
FFTs, a dense Jacobian buffer, FP32 dots) against the FP64 oracle, on the CPU.
The GPU tests at sizes the oracle cannot finish compare the library with this
restatement, so it is held here to the oracle first: transforms and
derivative columns to 1e-5 (complex64 arithmetic), three outer iterations of
JointRecon to 1e-4 on the image and objective, 1e-3 on motion, damping
identical."""
import numpy as np
import pytest
import torch

import bench
from oracle import motion_oracle as O

SHAPE = (16, 12, 8)


@pytest.fixture(scope="module")
def grids(me):
    kg, rkg = me.compute_transform_grids_voxel(SHAPE, (1, 1, 1), xp=torch)
    okg, orkg = O.compute_transform_grids_voxel(SHAPE, (1, 1, 1))
    return kg, rkg, okg, orkg


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_plain_transform_and_derivative_match_oracle(grids):
    kg, rkg, okg, orkg = grids
    rng = np.random.default_rng(0)
    x = rng.standard_normal(SHAPE) + 1j * rng.standard_normal(SHAPE)
    xt = torch.as_tensor(x.astype(np.complex64))
    th = np.array([0.7, -1.2, 0.4, 0.05, -0.08, 0.03])
    assert _rel(bench.plain_transform(torch, xt, th, kg, rkg).numpy(), O.rigid_apply(x, th, okg, orkg)) < 1e-5
    assert _rel(bench.plain_transform(torch, xt, th, kg, rkg, adjoint=True).numpy(),
                O.rigid_adjoint(x, th, okg, orkg)) < 1e-5
    for p in range(6):
        assert _rel(bench.plain_deriv(torch, xt, th, kg, rkg, p).numpy(), O.rigid_deriv(x, th, okg, orkg, p)) < 1e-5


def test_plain_joint_recon_matches_oracle(me, grids):
    kg, rkg, okg, orkg = grids
    sim = me.simulation
    gt = sim.phantom(SHAPE, seed=0).numpy().astype(np.complex128)
    mps = sim.coil_maps(SHAPE, 4, seed=1).numpy().astype(np.complex128)
    A, _ = sim.disorder_sampling(SHAPE, (4, 4), 2)
    t_true = sim.motion_uniform(A.shape[0], rot_deg=2.0, trans_vox=1.0, seed=2)
    ksp = O.generate_corrupted_kspace(gt, mps, A, t_true)
    ref = O.JointRecon(ksp, mps, A, okg, orkg, max_joint_iter=3)
    xr, tr = ref.run()
    pj = bench.PlainJointRecon(torch, torch.as_tensor(ksp), torch.as_tensor(mps), A, kg, rkg)
    obj = np.array([pj.step() for _ in range(3)])
    assert pj.max_norm == pytest.approx(ref.max_norm, rel=1e-5)
    assert _rel(pj.image().numpy(), xr) < 1e-4
    assert np.abs(pj.t - tr).max() < 1e-3
    assert np.max(np.abs(obj - ref.objective_history) / ref.objective_history) < 1e-4
    assert np.array_equal(pj.damp, ref.damp)
