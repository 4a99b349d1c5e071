"""Parity at the reference's CPU-runnable configuration itself: 64 x 64 x 64,
8 coils, 16 DISORDER shots (4 x 4 tiles, accel 1), known uniform motion,
30 dB noise (simulation.make_config("c1")).  The configuration's work is E,
E^H and CG-SENSE; the oracle's FP64 normal operator takes seconds here, so
the CG check runs its first step (right-hand side, initial residual, one
update) rather than all 20.  Relative L2 1e-4."""
import numpy as np
import pytest
import torch

from oracle import motion_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-4


def _rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm(a - np.asarray(b)) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def c1(me):
    d = me.simulation.make_config("c1", device=torch.device("cuda:0"))
    shape = tuple(d["config"]["shape"])
    ax = (-3, -2, -1)
    ksp_u = np.fft.ifftshift(d["ksp"].cpu().numpy().astype(np.complex128), axes=ax)
    mps_u = np.fft.ifftshift(d["mps"].cpu().numpy().astype(np.complex128), axes=ax)
    A_u = np.fft.ifftshift(d["A"], axes=(1, 2))
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    okg, orkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    ssq = np.sum(np.abs(mps_u) ** 2, axis=0)
    assert d["A"].shape[0] == 16 and mps_u.shape[0] == 8
    return dict(shape=shape, ksp_u=ksp_u, mps_u=mps_u, A_u=A_u, t=d["t_corr"], kg=kg, rkg=rkg, okg=okg,
                orkg=orkg, P=(1 / (ssq + 1e-3)).astype(np.float32), M=ssq > 0.1)


def test_c1_encoding_and_normal_operator(me, c1):
    d = c1
    pb = me.encoding.SenseProblem(d["ksp_u"], d["mps_u"], d["A_u"], d["kg"], d["rkg"])
    pb.theta.copy_(torch.as_tensor(d["t"]))
    rng = np.random.default_rng(9)
    x = rng.standard_normal(d["shape"]) + 1j * rng.standard_normal(d["shape"])
    Y = pb.forward(pb.image_to_plane(x)).cpu().numpy()
    pl, n2, nc = pb.plan, d["shape"][2], 8
    for s in range(pb.nshots):
        dense = O.encode_forward(x, d["mps_u"], d["A_u"][s], d["t"][s], d["okg"], d["orkg"])
        hyb = np.fft.ifft(dense, axis=-1, norm="ortho")
        o, c = pl.off_h[s], pl.cnt_h[s]
        ref = hyb[:, pl.k0_h[o:o + c], pl.k1_h[o:o + c], :]
        got = Y[o * n2 * nc:(o + c) * n2 * nc].reshape(n2, nc, c)
        assert _rel(np.transpose(got, (1, 2, 0)), ref) < TOL, s
    out = pb.zeros_image()
    pb.normal(pb.image_to_plane(x), out)
    ref = O.normal_op(x, d["mps_u"], d["A_u"], d["t"], d["okg"], d["orkg"], d["M"])
    assert _rel(pb.image_from_plane(out), ref) < TOL


def test_c1_cg_first_step(me, c1):
    d = c1
    args = (d["ksp_u"], d["mps_u"], d["A_u"], d["t"])
    got = me.estimate_image_cg(*args, d["kg"], d["rkg"], d["P"], d["M"], max_iter=1, tol=0.0)
    ref = O.estimate_image_cg(*args, d["okg"], d["orkg"], d["P"], d["M"], max_iter=1, tol=0.0)
    assert _rel(got, ref) < TOL


def test_c1_twenty_cg_steps_match_plain_restatement(me, c1):
    """The configuration's full work, 20 CG-SENSE steps with the known motion,
    against the plain torch.fft restatement of estimate_image_cg on the same
    GPU (bench.PlainJointRecon.cg; held to the oracle in
    tests/test_plain_reference.py): relative L2 1e-4."""
    import bench

    d = c1
    dev = torch.device("cuda:0")
    got = me.estimate_image_cg(d["ksp_u"], d["mps_u"], d["A_u"], d["t"], d["kg"], d["rkg"], d["P"], d["M"],
                               max_iter=20, tol=0.0)
    tkg, trkg = me.compute_transform_grids_voxel(d["shape"], (1, 1, 1), xp=torch)
    ax = (-3, -2, -1)
    pj = bench.PlainJointRecon(torch, torch.as_tensor(np.fft.fftshift(d["ksp_u"], axes=ax), device=dev),
                               torch.as_tensor(np.fft.fftshift(d["mps_u"], axes=ax), device=dev),
                               np.fft.fftshift(d["A_u"], axes=(1, 2)), tkg, trkg, t0=d["t"], max_cg_iter=20, tol=0.0)
    pj.cg()
    # PlainJointRecon normalises the data by max |ifftn(ksp)| as JointRecon does; CG is linear in the data
    assert _rel(got, (pj.x * pj.max_norm).cpu().numpy()) < TOL
