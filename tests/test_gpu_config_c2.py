"""Joint reconstruction at 128 x 128 x 96 with 16 coils and 32 DISORDER
shots (tile 8 x 8, acceleration 2), random-walk motion and 30 dB noise
(simulation.make_config("c2")), whose work is outer iterations of 10 CG
steps and one LM update of all 32 x 6 parameters.

* Against the FP64 oracle on two of the 32 shots (its dense FP64 path takes
  seconds per shot at this size): every sample of each shot's forward
  encoding, the normal operator over the two shots, one LM update (damping
  identical).
* All 32 shots: three outer iterations of JointRecon (10 CG + 1 LM each)
  against the plain torch.fft restatement of the reference on the same GPU
  (bench.PlainJointRecon, held to the oracle in tests/test_plain_reference.py).

Tolerances: relative L2 1e-4 on images, k-space and objectives; 1e-3
(degrees, voxels) on motion; LM damping identical."""
import os

import numpy as np
import pytest
import torch

from oracle import motion_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-4
MOTION_TOL = 1e-3
SUB = 2  # shots checked against the FP64 oracle


def _rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _motion_err(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    d = np.abs(a - np.asarray(b))
    return max(d[:, :3].max(), np.degrees(d[:, 3:]).max())


@pytest.fixture(scope="module")
def c2(me):
    O.set_workers(min(16, len(os.sched_getaffinity(0))))
    d = me.simulation.make_config("c2", device=torch.device("cuda:0"))
    cfg = d["config"]
    shape = tuple(cfg["shape"])
    assert shape == (128, 128, 96) and cfg["coils"] == 16 and d["A"].shape[0] == 32
    print(f"c2: {shape}, {cfg['coils']} coils, {d['A'].shape[0]} shots", flush=True)
    return dict(d, shape=shape)


@pytest.fixture(scope="module")
def c2_host(c2):
    ax = (-3, -2, -1)
    ksp_u = np.fft.ifftshift(c2["ksp"].cpu().numpy().astype(np.complex128), axes=ax)
    mps_u = np.fft.ifftshift(c2["mps"].cpu().numpy().astype(np.complex128), axes=ax)
    A_u = np.fft.ifftshift(c2["A"], axes=(1, 2))
    okg, orkg = O.compute_transform_grids_voxel(c2["shape"], (1, 1, 1))
    ssq = np.sum(np.abs(mps_u) ** 2, axis=0)
    return dict(ksp_u=ksp_u, mps_u=mps_u, A_u=A_u, okg=okg, orkg=orkg, M=ssq > 0.1)


def test_c2_subset_encoding_and_normal_operator(me, c2, c2_host):
    h = c2_host
    shape = c2["shape"]
    t = c2["t_corr"][:SUB]
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    pb = me.encoding.SenseProblem(h["ksp_u"], h["mps_u"], h["A_u"][:SUB], kg, rkg)
    pb.theta.copy_(torch.as_tensor(t))
    rng = np.random.default_rng(21)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    Y = pb.forward(pb.image_to_plane(x)).cpu().numpy()
    pl, n2, nc = pb.plan, shape[2], 16
    for s in range(SUB):
        dense = O.encode_forward(x, h["mps_u"], h["A_u"][s], t[s], h["okg"], h["orkg"])
        hyb = np.fft.ifft(dense, axis=-1, norm="ortho")
        o, c = pl.off_h[s], pl.cnt_h[s]
        ref = hyb[:, pl.k0_h[o:o + c], pl.k1_h[o:o + c], :]
        got = Y[o * n2 * nc:(o + c) * n2 * nc].reshape(n2, nc, c)
        assert _rel(np.transpose(got, (1, 2, 0)), ref) < TOL, s
    out = pb.zeros_image()
    pb.normal(pb.image_to_plane(x), out)
    ref = O.normal_op(x, h["mps_u"], h["A_u"][:SUB], t, h["okg"], h["orkg"], h["M"])
    assert _rel(pb.image_from_plane(out), ref) < TOL


def test_c2_subset_lm_update(me, c2, c2_host):
    h = c2_host
    shape = c2["shape"]
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    x = np.fft.ifftshift(c2["gt"].cpu().numpy().astype(np.complex128))
    t0 = c2["t_corr"][:SUB] + np.random.default_rng(22).normal(0, 0.05, (SUB, 6))
    t_ref, damp_ref, fl_ref = t0.copy(), np.ones(SUB), np.zeros(SUB, dtype=bool)
    x_ref = O.estimate_transform(h["ksp_u"], h["mps_u"], h["A_u"][:SUB], x, h["okg"], h["orkg"], damp_ref, fl_ref,
                                 t_ref)
    t_got, damp_got, fl_got = t0.copy(), np.ones(SUB), np.zeros(SUB, dtype=bool)
    x_got = me.estimate_transform(h["ksp_u"], h["mps_u"], h["A_u"][:SUB], x, kg, rkg, damp_got, fl_got, None, None,
                                  None, None, transforms=t_got)
    assert np.array_equal(damp_got, damp_ref)
    assert _motion_err(t_got, t_ref) < MOTION_TOL
    assert _rel(x_got, x_ref) < TOL


def test_c2_outer_iterations_match_plain_restatement(me, c2):
    import bench

    shape = c2["shape"]
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    n_outer = 3
    app = me.JointRecon(c2["ksp"], c2["mps"], c2["A"], kg, rkg, max_cg_iter=10, max_joint_iter=n_outer)
    x_got, t_got = app.run()
    tkg, trkg = me.compute_transform_grids_voxel(shape, (1, 1, 1), xp=torch)
    pj = bench.PlainJointRecon(torch, c2["ksp"], c2["mps"], c2["A"], tkg, trkg, max_cg_iter=10)
    obj = np.array([pj.step() for _ in range(n_outer)])
    print(f"c2: image rel {_rel(x_got, pj.image()):.2e}, motion {_motion_err(t_got, pj.t):.2e}, "
          f"accepted per iteration {[sum(a) for a in pj.accepted]}", flush=True)
    assert abs(app.max_norm - pj.max_norm) / pj.max_norm < 1e-5
    assert _rel(x_got, pj.image()) < TOL
    assert _motion_err(t_got, pj.t) < MOTION_TOL
    assert np.array_equal(app.damp.cpu().numpy(), pj.damp)
    assert np.allclose(app.objective_history, obj, rtol=TOL, atol=0)
    assert obj[-1] < obj[0]
