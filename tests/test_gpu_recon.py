"""GPU parity of encoding, CG-SENSE, LM and JointRecon against the FP64 oracle.

Tolerances (written here, from the build's target): relative L2 1e-4 on
images, k-space and residuals; 1e-3 (degrees for rotations, voxels for
translations) on estimated motion.  The problem is small enough for the dense
oracle (32 x 24 x 16, 4 coils, 8 DISORDER shots) and uses seeded synthetic
inputs; the data are motion-corrupted with known per-shot parameters plus
noise."""
import numpy as np
import pytest
import torch

from oracle import motion_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-4
MOTION_TOL = 1e-3

SHAPE = (32, 24, 16)
NC = 4


def _rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _motion_err(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    d = np.abs(np.asarray(a) - np.asarray(b))
    return max(d[:, :3].max(), np.degrees(d[:, 3:]).max())


@pytest.fixture(scope="module")
def prob_data(me):
    sim = me.simulation
    gt = sim.phantom(SHAPE, seed=0).numpy().astype(np.complex128)
    mps = sim.coil_maps(SHAPE, NC, seed=1).numpy().astype(np.complex128)
    A, order = sim.disorder_sampling(SHAPE, (4, 4), 2)
    t_true = sim.motion_uniform(A.shape[0], rot_deg=3.0, trans_vox=1.5, seed=2)
    ksp = O.generate_corrupted_kspace(gt, mps, A, t_true)
    rng = np.random.default_rng(3)
    ksp = ksp + 1e-3 * np.abs(ksp).max() * (rng.standard_normal(ksp.shape) + 1j * rng.standard_normal(ksp.shape))
    # uncentred copies (the frame estimate_image_cg / estimate_transform expect)
    ax = (-3, -2, -1)
    ksp_u = np.fft.ifftshift(ksp, axes=ax)
    mps_u = np.fft.ifftshift(mps, axes=ax)
    A_u = np.fft.ifftshift(A, axes=(1, 2))
    kg, rkg = me.compute_transform_grids_voxel(SHAPE, (1, 1, 1))
    okg, orkg = O.compute_transform_grids_voxel(SHAPE, (1, 1, 1))
    ssq = np.sum(np.abs(mps_u) ** 2, axis=0)
    P = (1 / (ssq + 1e-3)).astype(np.float32)
    M = ssq > 0.1
    return dict(gt=gt, mps=mps, A=A, ksp=ksp, t_true=t_true, ksp_u=ksp_u, mps_u=mps_u, A_u=A_u, kg=kg, rkg=rkg,
                okg=okg, orkg=orkg, P=P, M=M)


def _problem(me, d):
    return me.encoding.SenseProblem(d["ksp_u"], d["mps_u"], d["A_u"], d["kg"], d["rkg"])


def test_corrupted_kspace_matches_oracle(me, prob_data):
    d = prob_data
    got = me.simulation.generate_corrupted_kspace(d["gt"], d["mps"], d["A"], d["t_true"])
    ref = O.generate_corrupted_kspace(d["gt"], d["mps"], d["A"], d["t_true"])
    assert _rel(got, ref) < TOL


def test_forward_encoding_matches_oracle(me, prob_data):
    d = prob_data
    pb = _problem(me, d)
    pb.theta.copy_(torch.as_tensor(d["t_true"]))
    rng = np.random.default_rng(4)
    x = rng.standard_normal(SHAPE) + 1j * rng.standard_normal(SHAPE)
    Y = pb.forward(pb.image_to_plane(x)).cpu().numpy()
    pl = pb.plan
    n2 = SHAPE[2]
    for s in range(pb.nshots):
        dense = O.encode_forward(x, d["mps_u"], d["A_u"][s], d["t_true"][s], d["okg"], d["orkg"])
        hyb = np.fft.ifft(dense, axis=-1, norm="ortho")  # readout inverse FFT -> hybrid space
        o, c = pl.off_h[s], pl.cnt_h[s]
        ref = hyb[:, pl.k0_h[o:o + c], pl.k1_h[o:o + c], :]  # (C, cnt, n2)
        got = Y[o * n2 * NC:(o + c) * n2 * NC].reshape(n2, NC, c)
        assert _rel(np.transpose(got, (1, 2, 0)), ref) < TOL
        # nothing outside the mask
        assert np.abs(dense[:, ~d["A_u"][s][..., 0]]).max() == 0


def test_encoding_adjoint_identity(me, prob_data):
    d = prob_data
    pb = _problem(me, d)
    rng = np.random.default_rng(5)
    x = pb.image_to_plane(rng.standard_normal(SHAPE) + 1j * rng.standard_normal(SHAPE))
    U = torch.stack([x] * pb.nshots)
    Y = torch.empty_like(pb.ymeas)
    pb._encode(U, pb.vol, 0, pb.nshots, y=Y)
    y2 = torch.randn_like(Y)
    W = torch.empty((pb.nshots * pb.ncg, pb.vol), dtype=torch.complex64, device=U.device)
    pb._decode(y2, 0, pb.nshots, W)
    W = W.reshape(pb.nshots, pb.ncg, pb.vol)[:, :pb.parts].sum(1)  # parts 0 and 1 hold the sum
    lhs = torch.vdot(Y, y2).cpu().numpy()
    rhs = sum(torch.vdot(U[s], W[s]).cpu().numpy() for s in range(pb.nshots))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_normal_op_and_rhs_match_oracle(me, prob_data):
    d = prob_data
    pb = _problem(me, d)
    pb.theta.copy_(torch.as_tensor(d["t_true"]))
    rng = np.random.default_rng(6)
    x = rng.standard_normal(SHAPE) + 1j * rng.standard_normal(SHAPE)
    out = pb.zeros_image()
    pb.normal(pb.image_to_plane(x), out)
    ref = O.normal_op(x, d["mps_u"], d["A_u"], d["t_true"], d["okg"], d["orkg"], d["M"])
    assert _rel(pb.image_from_plane(out), ref) < TOL
    b = pb.rhs(cache=False)
    refb = O.cg_rhs(d["ksp_u"], d["mps_u"], d["A_u"], d["t_true"], d["okg"], d["orkg"], d["M"])
    assert _rel(pb.image_from_plane(b), refb) < TOL
    b2 = pb.rhs(cache=True)
    assert _rel(pb.image_from_plane(b2), refb) < TOL


def test_objective_matches_oracle(me, prob_data):
    d = prob_data
    pb = _problem(me, d)
    pb.theta.copy_(torch.as_tensor(d["t_true"]))
    x = d["gt"] * 0.9
    got = pb.objective(pb.image_to_plane(np.fft.ifftshift(x)))
    ref = sum(np.linalg.norm(O.shot_residual(d["ksp_u"], d["mps_u"], d["A_u"][s], np.fft.ifftshift(x),
                                             d["t_true"][s], d["okg"], d["orkg"])) ** 2 for s in range(8))
    assert abs(got - ref) / ref < TOL


@pytest.mark.parametrize("max_iter,tol", [(5, 1e-6), (20, 1e-2)])
def test_estimate_image_cg_matches_oracle(me, prob_data, max_iter, tol):
    d = prob_data
    args = (d["ksp_u"], d["mps_u"], d["A_u"], d["t_true"])
    got = me.estimate_image_cg(*args, d["kg"], d["rkg"], d["P"], d["M"], max_iter=max_iter, tol=tol)
    ref = O.estimate_image_cg(*args, d["okg"], d["orkg"], d["P"], d["M"], max_iter=max_iter, tol=tol)
    assert _rel(got, ref) < TOL


def test_estimate_transform_matches_oracle(me, prob_data):
    d = prob_data
    x = O.estimate_image_cg(d["ksp_u"], d["mps_u"], d["A_u"], d["t_true"], d["okg"], d["orkg"], d["P"], d["M"],
                            max_iter=5)
    t0 = d["t_true"] + np.random.default_rng(7).normal(0, 0.05, d["t_true"].shape)
    ns = t0.shape[0]
    t_ref, damp_ref, fl_ref = t0.copy(), np.ones(ns), np.zeros(ns, dtype=bool)
    trace = []
    x_ref = O.estimate_transform(d["ksp_u"], d["mps_u"], d["A_u"], x, d["okg"], d["orkg"], damp_ref, fl_ref, t_ref,
                                 trace=trace)
    t_got, damp_got, fl_got = t0.copy(), np.ones(ns), np.zeros(ns, dtype=bool)
    x_got = me.estimate_transform(d["ksp_u"], d["mps_u"], d["A_u"], x, d["kg"], d["rkg"], damp_got, fl_got, None, None,
                                  None, None, transforms=t_got)
    assert np.array_equal(damp_got, damp_ref), (damp_got, damp_ref, [tr[:3] for tr in trace])
    assert _motion_err(t_got, t_ref) < MOTION_TOL
    assert _rel(x_got, x_ref) < TOL


def test_estimate_transform_geodesic_acceleration_matches_oracle(me, prob_data):
    """use_ga=True (transform_estimation.py:73-121) against the oracle's FP64
    forward difference at eps = 1e-4.  The library estimates J'v by an FP32
    central difference at h = 1e-3 (an FP32 forward difference at 1e-4, the
    reference's own arithmetic, deviates from the FP64 one by about 8e-4 deg
    here); the plain step's bars hold: motion 1e-3 (deg / voxel), image 1e-4,
    damping identical, and the acceleration itself must be visible (its
    effect on the motion is 6e-2 here)."""
    d = prob_data
    x = O.estimate_image_cg(d["ksp_u"], d["mps_u"], d["A_u"], d["t_true"], d["okg"], d["orkg"], d["P"], d["M"],
                            max_iter=5)
    t0 = d["t_true"] + np.random.default_rng(8).normal(0, 0.05, d["t_true"].shape)
    ns = t0.shape[0]
    t_ref, damp_ref, fl_ref = t0.copy(), np.ones(ns), np.zeros(ns, dtype=bool)
    x_ref = O.estimate_transform(d["ksp_u"], d["mps_u"], d["A_u"], x, d["okg"], d["orkg"], damp_ref, fl_ref, t_ref,
                                 use_ga=True)
    t_plain = t0.copy()
    O.estimate_transform(d["ksp_u"], d["mps_u"], d["A_u"], x, d["okg"], d["orkg"], np.ones(ns),
                         np.zeros(ns, dtype=bool), t_plain)
    effect = _motion_err(t_plain, t_ref)  # what the acceleration changes
    t_got, damp_got, fl_got = t0.copy(), np.ones(ns), np.zeros(ns, dtype=bool)
    x_got = me.estimate_transform(d["ksp_u"], d["mps_u"], d["A_u"], x, d["kg"], d["rkg"], damp_got, fl_got, None, None,
                                  None, None, transforms=t_got, use_ga=True)
    assert np.array_equal(damp_got, damp_ref)
    err = _motion_err(t_got, t_ref)
    print(f"GA: effect {effect:.3e}, GPU vs oracle {err:.3e}")
    assert effect > 10 * MOTION_TOL
    assert err < MOTION_TOL
    assert _rel(x_got, x_ref) < TOL


def test_estimate_transform_requires_transforms(me, prob_data):
    d = prob_data
    with pytest.raises(TypeError):
        me.estimate_transform(d["ksp_u"], d["mps_u"], d["A_u"], d["gt"], d["kg"], d["rkg"], np.ones(8),
                              np.zeros(8, bool), None, None, None, None, transforms=None)


def test_joint_recon_matches_oracle(me, prob_data):
    d = prob_data
    kw = dict(max_cg_iter=3, max_joint_iter=3)
    app = me.JointRecon(d["ksp"], d["mps"], d["A"], d["kg"], d["rkg"], **kw)
    x_got, t_got = app.run()
    ref = O.JointRecon(d["ksp"], d["mps"], d["A"], d["okg"], d["orkg"], **kw)
    x_ref, t_ref = ref.run()
    assert abs(app.max_norm - ref.max_norm) / ref.max_norm < 1e-5
    assert _rel(x_got, x_ref) < TOL
    assert _motion_err(t_got, t_ref) < MOTION_TOL
    assert np.allclose(app.objective_history, ref.objective_history, rtol=TOL, atol=0)
    assert app.transform_history.shape == ref.transform_history.shape


@pytest.mark.parametrize("ncoils", [20, 40])
def test_normal_op_many_coil_groups(me, ncoils):
    """Coil counts that are not multiples of the 16-coil forward groups, and
    more than two groups."""
    shape = (16, 16, 8)
    sim = me.simulation
    mps = sim.coil_maps(shape, ncoils, seed=4).numpy().astype(np.complex128)
    A, _ = sim.disorder_sampling(shape, (4, 4), 2)
    t = sim.motion_uniform(A.shape[0], rot_deg=3.0, trans_vox=1.0, seed=5)
    rng = np.random.default_rng(8)
    ksp = rng.standard_normal((ncoils, *shape)) + 1j * rng.standard_normal((ncoils, *shape))
    A_u = np.fft.ifftshift(A, axes=(1, 2))
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    okg, orkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    pb = me.encoding.SenseProblem(ksp, mps, A_u, kg, rkg)
    pb.theta.copy_(torch.as_tensor(t))
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    out = pb.zeros_image()
    pb.normal(pb.image_to_plane(x), out)
    M = np.sum(np.abs(mps) ** 2, axis=0) > 0.1
    ref = O.normal_op(x, mps, A_u, t, okg, orkg, M)
    assert _rel(pb.image_from_plane(out), ref) < TOL
    refb = O.cg_rhs(ksp, mps, A_u, t, okg, orkg, M)
    assert _rel(pb.image_from_plane(pb.rhs(cache=True)), refb) < TOL
    assert _rel(pb.image_from_plane(pb.rhs(cache=False)), refb) < TOL


def test_chain_derivative_columns_match_oracle(me, prob_data):
    """Jacobian images from the shared rotation chain (columns 0-5 in the
    order the LM step uses them) against the oracle's derivative."""
    d = prob_data
    pb = _problem(me, d)
    th = torch.as_tensor(d["t_true"], device=pb.device)
    rng = np.random.default_rng(9)
    x = rng.standard_normal(SHAPE) + 1j * rng.standard_normal(SHAPE)
    xp = pb.image_to_plane(x)
    U = torch.empty((pb.nshots, pb.vol), dtype=torch.complex64, device=pb.device)
    for p_idx in range(6):
        pb._deriv_chain(xp, U, th, p_idx, p_idx == 0)
        for s in (0, 3, 7):
            ref = O.rigid_deriv(x, d["t_true"][s], d["okg"], d["orkg"], p_idx)
            assert _rel(pb.image_from_plane(U[s]), ref) < TOL, (p_idx, s)


def test_cg_tracking_and_initial_norms(me, prob_data):
    """E T x carried through CG by alpha E T p equals a fresh encoding of the
    final x; the first encoding's residual norms equal the objective."""
    d = prob_data
    pb = _problem(me, d)
    pb.theta.copy_(torch.as_tensor(d["t_true"]))
    rng = np.random.default_rng(10)
    x = pb.image_to_plane(0.1 * (rng.standard_normal(SHAPE) + 1j * rng.standard_normal(SHAPE)))
    obj0 = pb.objective_per_shot(x).cpu().numpy()
    b = pb.rhs()
    track = pb.track_buffer()
    norms0 = torch.empty(pb.nshots, dtype=torch.float64, device=pb.device)
    pb.cg(x, b, 4, 1e-9, check_every=0, track=track, norms0=norms0)
    np.testing.assert_allclose(norms0.cpu().numpy(), obj0, rtol=1e-6)
    fresh = pb.forward(x).clone()
    assert _rel(track, fresh.cpu().numpy()) < TOL


def test_cg_residual_form_matches_explicit_rhs(me, prob_data):
    """cg(b=None) forms r0 = b - A x0 from one normal operator of
    E T x0 - y: the iterates equal those with the explicit right-hand side,
    the tracked buffer ends as the residual E T x - y of the final x, and
    estimate_image_cg (which uses it) matches the oracle."""
    d = prob_data
    rng = np.random.default_rng(13)
    x0 = 0.1 * (rng.standard_normal(SHAPE) + 1j * rng.standard_normal(SHAPE))
    pb = _problem(me, d)
    pb.theta.copy_(torch.as_tensor(d["t_true"]))
    xa = pb.image_to_plane(x0)
    pb.cg(xa, pb.rhs(), 4, 1e-12, check_every=0)
    pb2 = _problem(me, d)
    pb2.theta.copy_(torch.as_tensor(d["t_true"]))
    xb = pb2.image_to_plane(x0)
    track = pb2.track_buffer()
    norms0 = torch.empty(pb2.nshots, dtype=torch.float64, device=pb2.device)
    pb2.cg(xb, None, 4, 1e-12, check_every=0, track=track, norms0=norms0)
    assert _rel(xb, xa.cpu().numpy()) < 1e-5
    np.testing.assert_allclose(norms0.cpu().numpy(), pb2.objective_per_shot(pb2.image_to_plane(x0)).cpu().numpy(),
                               rtol=1e-6)
    fresh = pb2.forward(xb, meas=True).clone()
    assert _rel(track, fresh.cpu().numpy()) < TOL
    got = me.estimate_image_cg(d["ksp_u"], d["mps_u"], d["A_u"], d["t_true"], d["kg"], d["rkg"], d["P"], d["M"],
                               x0=x0, max_iter=5)
    ref = O.estimate_image_cg(d["ksp_u"], d["mps_u"], d["A_u"], d["t_true"], d["okg"], d["orkg"], d["P"], d["M"],
                              x0=x0, max_iter=5)
    assert _rel(got, ref) < TOL


def test_empty_shot_matches_oracle(me, prob_data):
    """A shot with no samples (a user mask can hold one): it contributes
    nothing to the normal operator and right-hand side, and its LM step is
    rejected (e1 = e0 = 0), so its damping grows by 1.5 as in the reference."""
    d = prob_data
    A_u = d["A_u"].copy()
    A_u[3] = False
    pb = me.encoding.SenseProblem(d["ksp_u"], d["mps_u"], A_u, d["kg"], d["rkg"])
    assert int(pb.plan.cnt_h[3]) == 0
    pb.theta.copy_(torch.as_tensor(d["t_true"]))
    rng = np.random.default_rng(12)
    x = rng.standard_normal(SHAPE) + 1j * rng.standard_normal(SHAPE)
    out = pb.zeros_image()
    pb.normal(pb.image_to_plane(x), out)
    ref = O.normal_op(x, d["mps_u"], A_u, d["t_true"], d["okg"], d["orkg"], d["M"])
    assert _rel(pb.image_from_plane(out), ref) < TOL
    refb = O.cg_rhs(d["ksp_u"], d["mps_u"], A_u, d["t_true"], d["okg"], d["orkg"], d["M"])
    assert _rel(pb.image_from_plane(pb.rhs(cache=False)), refb) < TOL
    xi = O.estimate_image_cg(d["ksp_u"], d["mps_u"], A_u, d["t_true"], d["okg"], d["orkg"], d["P"], d["M"], max_iter=5)
    t0 = d["t_true"] + np.random.default_rng(7).normal(0, 0.05, d["t_true"].shape)
    ns = t0.shape[0]
    t_ref, damp_ref, fl_ref = t0.copy(), np.ones(ns), np.zeros(ns, dtype=bool)
    x_ref = O.estimate_transform(d["ksp_u"], d["mps_u"], A_u, xi, d["okg"], d["orkg"], damp_ref, fl_ref, t_ref)
    t_got, damp_got, fl_got = t0.copy(), np.ones(ns), np.zeros(ns, dtype=bool)
    x_got = me.estimate_transform(d["ksp_u"], d["mps_u"], A_u, xi, d["kg"], d["rkg"], damp_got, fl_got, None, None,
                                  None, None, transforms=t_got)
    assert damp_ref[3] == 1.5 and np.array_equal(damp_got, damp_ref)
    assert _motion_err(t_got, t_ref) < MOTION_TOL
    assert _rel(x_got, x_ref) < TOL


def test_masks_along_the_readout_axis(me, prob_data):
    """Quirk Q16: the readout is axis 2 and the masks are singleton there.  A
    mask repeated along axis 2 is the same mask; one that varies along it
    cannot be encoded in hybrid space and is refused."""
    d = prob_data
    rep = np.repeat(d["A_u"], SHAPE[2], axis=3)
    a = me.encoding.ShotPlan(d["A_u"], torch.device("cuda:0"))
    b = me.encoding.ShotPlan(rep, torch.device("cuda:0"))
    assert np.array_equal(a.cnt_h, b.cnt_h) and np.array_equal(a.k0_h, b.k0_h) and np.array_equal(a.k1_h, b.k1_h)
    rep[0, 0, 0, 1] = not rep[0, 0, 0, 0]
    with pytest.raises(ValueError, match="readout axis"):
        me.encoding.ShotPlan(rep, torch.device("cuda:0"))


def test_fractional_support_mask_is_refused(me, prob_data):
    """The kernels apply M as a 0/1 support mask; the reference's M is the
    boolean sum|S|^2 > 0.1 (joint_recon.py:85), so a fractional M is an
    error rather than being silently truncated."""
    d = prob_data
    args = (d["ksp_u"], d["mps_u"], d["A_u"], d["t_true"], d["kg"], d["rkg"], d["P"])
    with pytest.raises(ValueError):
        me.estimate_image_cg(*args, 0.5 * d["M"].astype(np.float32), max_iter=1)
    got = me.estimate_image_cg(*args, d["M"].astype(np.float32), max_iter=1)  # 0/1 floats are fine
    ref = me.estimate_image_cg(*args, d["M"], max_iter=1)
    assert torch.equal(got, ref)
