"""The shot-batched paths against the FP64 oracle.  A problem whose shots do
not fit one working set is processed in batches (accumulated adjoint sums,
cached right-hand side per batch, CG with tracked encodings, LM and the
objective per batch).  The default budget keeps every configuration here in
one batch, so this module lowers it to 3 shots per batch (8 shots: 3 + 3 + 2)
and repeats the main parity checks.  Tolerances as in test_gpu_recon."""
import numpy as np
import pytest
import torch

from oracle import motion_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-4
MOTION_TOL = 1e-3
SHAPE = (32, 24, 16)
NC = 4
VOL = SHAPE[0] * SHAPE[1] * SHAPE[2]


def _rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm(a - np.asarray(b)) / np.linalg.norm(b))


def _motion_err(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    d = np.abs(a - np.asarray(b))
    return max(d[:, :3].max(), np.degrees(d[:, 3:]).max())


@pytest.fixture(scope="module")
def small_batches(me):
    old = me.encoding.DEFAULT_MEM_BUDGET
    me.encoding.DEFAULT_MEM_BUDGET = VOL * 8 * 3  # 3 shot volumes per batch
    yield
    me.encoding.DEFAULT_MEM_BUDGET = old


@pytest.fixture(scope="module")
def data(me, small_batches):
    sim = me.simulation
    gt = sim.phantom(SHAPE, seed=0).numpy().astype(np.complex128)
    mps = sim.coil_maps(SHAPE, NC, seed=1).numpy().astype(np.complex128)
    A, _ = sim.disorder_sampling(SHAPE, (4, 4), 2)
    t_true = sim.motion_uniform(A.shape[0], rot_deg=3.0, trans_vox=1.5, seed=2)
    ksp = O.generate_corrupted_kspace(gt, mps, A, t_true)
    rng = np.random.default_rng(3)
    ksp = ksp + 1e-3 * np.abs(ksp).max() * (rng.standard_normal(ksp.shape) + 1j * rng.standard_normal(ksp.shape))
    ax = (-3, -2, -1)
    ksp_u, mps_u = np.fft.ifftshift(ksp, axes=ax), np.fft.ifftshift(mps, axes=ax)
    A_u = np.fft.ifftshift(A, axes=(1, 2))
    kg, rkg = me.compute_transform_grids_voxel(SHAPE, (1, 1, 1))
    okg, orkg = O.compute_transform_grids_voxel(SHAPE, (1, 1, 1))
    ssq = np.sum(np.abs(mps_u) ** 2, axis=0)
    return dict(gt=gt, mps=mps, A=A, ksp=ksp, t=t_true, ksp_u=ksp_u, mps_u=mps_u, A_u=A_u, kg=kg, rkg=rkg, okg=okg,
                orkg=orkg, P=(1 / (ssq + 1e-3)).astype(np.float32), M=ssq > 0.1)


def test_normal_operator_and_rhs(me, data):
    d = data
    pb = me.encoding.SenseProblem(d["ksp_u"], d["mps_u"], d["A_u"], d["kg"], d["rkg"])
    assert pb.batch == 3 and len(list(pb.batches())) == 3
    pb.theta.copy_(torch.as_tensor(d["t"]))
    x = np.random.default_rng(6).standard_normal(SHAPE) + 1j * np.random.default_rng(7).standard_normal(SHAPE)
    out = pb.zeros_image()
    pb.normal(pb.image_to_plane(x), out)
    assert _rel(pb.image_from_plane(out), O.normal_op(x, d["mps_u"], d["A_u"], d["t"], d["okg"], d["orkg"],
                                                      d["M"])) < TOL
    refb = O.cg_rhs(d["ksp_u"], d["mps_u"], d["A_u"], d["t"], d["okg"], d["orkg"], d["M"])
    assert _rel(pb.image_from_plane(pb.rhs(cache=False)), refb) < TOL
    assert _rel(pb.image_from_plane(pb.rhs(cache=True)), refb) < TOL
    assert _rel(pb.image_from_plane(pb.rhs(cache=True)), refb) < TOL  # from the per-batch cache


def test_estimate_image_cg(me, data):
    d = data
    args = (d["ksp_u"], d["mps_u"], d["A_u"], d["t"])
    got = me.estimate_image_cg(*args, d["kg"], d["rkg"], d["P"], d["M"], max_iter=5, tol=1e-6)
    ref = O.estimate_image_cg(*args, d["okg"], d["orkg"], d["P"], d["M"], max_iter=5, tol=1e-6)
    assert _rel(got, ref) < TOL


def test_estimate_transform(me, data):
    d = data
    x = O.estimate_image_cg(d["ksp_u"], d["mps_u"], d["A_u"], d["t"], d["okg"], d["orkg"], d["P"], d["M"],
                            max_iter=5)
    t0 = d["t"] + np.random.default_rng(7).normal(0, 0.05, d["t"].shape)
    ns = t0.shape[0]
    t_ref, damp_ref, fl_ref = t0.copy(), np.ones(ns), np.zeros(ns, dtype=bool)
    x_ref = O.estimate_transform(d["ksp_u"], d["mps_u"], d["A_u"], x, d["okg"], d["orkg"], damp_ref, fl_ref, t_ref)
    t_got, damp_got, fl_got = t0.copy(), np.ones(ns), np.zeros(ns, dtype=bool)
    x_got = me.estimate_transform(d["ksp_u"], d["mps_u"], d["A_u"], x, d["kg"], d["rkg"], damp_got, fl_got, None,
                                  None, None, None, transforms=t_got)
    assert np.array_equal(damp_got, damp_ref)
    assert _motion_err(t_got, t_ref) < MOTION_TOL
    assert _rel(x_got, x_ref) < TOL


def test_estimate_transform_geodesic_acceleration(me, data):
    """use_ga=True over three shot batches (the central-difference image is a
    per-batch buffer): same bars as the plain step."""
    d = data
    x = O.estimate_image_cg(d["ksp_u"], d["mps_u"], d["A_u"], d["t"], d["okg"], d["orkg"], d["P"], d["M"],
                            max_iter=5)
    t0 = d["t"] + np.random.default_rng(9).normal(0, 0.05, d["t"].shape)
    ns = t0.shape[0]
    t_ref, damp_ref = t0.copy(), np.ones(ns)
    x_ref = O.estimate_transform(d["ksp_u"], d["mps_u"], d["A_u"], x, d["okg"], d["orkg"], damp_ref,
                                 np.zeros(ns, dtype=bool), t_ref, use_ga=True)
    t_got, damp_got = t0.copy(), np.ones(ns)
    x_got = me.estimate_transform(d["ksp_u"], d["mps_u"], d["A_u"], x, d["kg"], d["rkg"], damp_got,
                                  np.zeros(ns, dtype=bool), None, None, None, None, transforms=t_got, use_ga=True)
    assert np.array_equal(damp_got, damp_ref)
    assert _motion_err(t_got, t_ref) < MOTION_TOL
    assert _rel(x_got, x_ref) < TOL


def test_joint_recon(me, data):
    d = data
    kw = dict(max_cg_iter=3, max_joint_iter=3)
    app = me.JointRecon(d["ksp"], d["mps"], d["A"], d["kg"], d["rkg"], **kw)
    assert app.problem.batch == 3
    x_got, t_got = app.run()
    ref = O.JointRecon(d["ksp"], d["mps"], d["A"], d["okg"], d["orkg"], **kw)
    x_ref, t_ref = ref.run()
    assert _rel(x_got, x_ref) < TOL
    assert _motion_err(t_got, t_ref) < MOTION_TOL
    assert np.allclose(app.objective_history, ref.objective_history, rtol=TOL, atol=0)


def test_data_consistency_gradients(me, data):
    d = data
    pb = me.encoding.SenseProblem(d["ksp_u"], d["mps_u"], d["A_u"], d["kg"], d["rkg"])
    x0 = d["gt"] * 0.8
    x = torch.as_tensor(np.fft.ifftshift(x0).astype(np.complex64), device=pb.device).requires_grad_(True)
    th = torch.as_tensor(d["t"], device=pb.device).requires_grad_(True)
    loss = me.data_consistency_loss(x, th, pb)
    loss.backward()
    xu = np.fft.ifftshift(x0)
    ref = O.dc_loss(xu, d["t"], d["ksp_u"], d["mps_u"], d["A_u"], d["okg"], d["orkg"])
    gx, gt = O.dc_loss_grad(xu, d["t"], d["ksp_u"], d["mps_u"], d["A_u"], d["okg"], d["orkg"])
    assert abs(float(loss.detach()) - ref) / ref < TOL
    assert _rel(x.grad, gx) < TOL
    assert _rel(th.grad, gt) < TOL


def test_sense_recon_one_pseudo_shot_per_batch(me, data):
    d = data
    old = me.encoding.DEFAULT_MEM_BUDGET
    me.encoding.DEFAULT_MEM_BUDGET = VOL * 8  # one pseudo-shot (512 lines) per batch
    try:
        y = d["ksp"]  # all 768 lines: two pseudo-shots of at most 512 lines, one per batch
        rec = me.SenseRecon(y, d["mps"], max_iter=10)
        assert rec.pb.nshots == 2 and len(list(rec.pb.batches())) == 2
        got = rec.run()
    finally:
        me.encoding.DEFAULT_MEM_BUDGET = old
    ref = O.sense_recon(y, d["mps"], max_iter=10)
    assert _rel(got, ref) < TOL
