"""GPU checks of the data-consistency loss and its hand-written backward.

* Jacobian columns of the apply order (me_rigid_deriv_apply) against the FP64
  oracle ``rigid_deriv_apply`` (itself pinned by central differences of
  ``rigid_apply``): relative L2 1e-4.
* Loss and both gradients against the FP64 oracle ``dc_loss`` /
  ``dc_loss_grad``: relative 1e-4 (loss, image gradient, parameter gradient).
* Central finite differences of the GPU loss itself at step 1e-3 along 64
  random (image, motion) directions: median relative error below 1e-3, and
  every error below 1e-3 of |g_x||dx| + |g_theta||dtheta| (the difference
  quotient of FP32 data is noisy for directions nearly orthogonal to the
  gradient, where the plain relative error is not meaningful).
* recompute=True and recompute=False give identical gradients.
* The columns' dots formed inside their last line pass equal the stored
  columns' FP64 dots (to FP32 rounding of the columns); the forward
  transform built through the chain equals the plain apply bit for bit.
"""
import numpy as np
import pytest
import torch

from oracle import motion_oracle as O

pytestmark = pytest.mark.gpu

SHAPE = (16, 12, 8)
NC = 4


def _rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def dc(me):
    sim = me.simulation
    rng = np.random.default_rng(40)
    mps = sim.coil_maps(SHAPE, NC, seed=41).numpy().astype(np.complex128)
    A, _ = sim.disorder_sampling(SHAPE, (4, 4), 2)
    A_u = np.fft.ifftshift(A, axes=(1, 2))
    mps_u = np.fft.ifftshift(mps, axes=(-3, -2, -1))
    ksp = rng.standard_normal((NC, *SHAPE)) + 1j * rng.standard_normal((NC, *SHAPE))
    t = sim.motion_uniform(A.shape[0], rot_deg=4.0, trans_vox=1.5, seed=42)
    x = rng.standard_normal(SHAPE) + 1j * rng.standard_normal(SHAPE)
    kg, rkg = me.compute_transform_grids_voxel(SHAPE, (1, 1, 1))
    okg, orkg = O.compute_transform_grids_voxel(SHAPE, (1, 1, 1))
    pb = me.encoding.SenseProblem(ksp, mps_u, A_u, kg, rkg)
    return dict(mps=mps_u, A=A_u, ksp=ksp, t=t, x=x, okg=okg, orkg=orkg, pb=pb)


def test_apply_order_columns_match_oracle(me, dc):
    pb = dc["pb"]
    th = torch.as_tensor(dc["t"], device=pb.device)
    xp = pb.image_to_plane(dc["x"])
    B = pb.nshots
    vol = pb.vol
    bufs = torch.empty((5, B, vol), dtype=torch.complex64, device=pb.device)
    u, c1, c2, col, work = bufs
    pb._rigid(xp, 0, u, th, 0)
    for p in (0, 1, 2, 3, 4, 5):
        me.autograd._deriv_apply(pb, xp, u, vol, c1, c2, col, work, th, p, p == 3)
        for s in (0, 5):
            ref = O.rigid_deriv_apply(dc["x"], dc["t"][s], dc["okg"], dc["orkg"], p)
            assert _rel(pb.image_from_plane(col[s]), ref) < 1e-4, (p, s)


def test_apply_chain_equals_apply(me, dc):
    """me_rigid_apply_chain's u is me_rigid_transform's apply bit for bit
    (the same passes in the same order), and its c1, c2 are the chain the
    columns start from."""
    pb = dc["pb"]
    th = torch.as_tensor(dc["t"], device=pb.device)
    xp = pb.image_to_plane(dc["x"])
    B, vol = pb.nshots, pb.vol
    bufs = torch.empty((7, B, vol), dtype=torch.complex64, device=pb.device)
    ref, u, c1, c2, d1, d2, work = bufs
    pb._rigid(xp, 0, ref, th, 0)
    me.autograd._apply_chain(pb, xp, c1, c2, u, th)
    assert torch.equal(u, ref)
    col = torch.empty_like(u)
    me.autograd._deriv_apply(pb, xp, u, vol, d1, d2, col, work, th, 3, True)  # builds its own chain
    assert torch.equal(d1, c1) and torch.equal(d2, c2)


def test_column_dots_without_storing_the_columns(me, dc):
    """me_rigid_deriv_apply_dot (the column's last pass reduces against w)
    equals storing the column and taking the FP64 batched dot, for every p.
    The two last passes are separately compiled kernels whose FP32 operation
    sequences (fused multiply-adds) may differ, so the columns agree to FP32
    rounding and their dots to about 1e-7 of the largest."""
    pb = dc["pb"]
    th = torch.as_tensor(dc["t"], device=pb.device)
    xp = pb.image_to_plane(dc["x"])
    B, vol = pb.nshots, pb.vol
    bufs = torch.empty((6, B, vol), dtype=torch.complex64, device=pb.device)
    u, c1, c2, col, work, w = bufs
    w.copy_(torch.randn((B, vol), dtype=torch.complex64, device=pb.device))
    pb._rigid(xp, 0, u, th, 0)
    fused = torch.zeros((B, 6), dtype=torch.float64, device=pb.device)
    plain = torch.zeros((B, 6), dtype=torch.float64, device=pb.device)
    for p in range(6):
        me.autograd._deriv_apply(pb, xp, u, vol, c1, c2, col, work, th, p, p == 3)
        me.autograd._dots(pb, w, col, B, plain[:, p:], 6)
        me.autograd._deriv_apply_dot(pb, xp, u, vol, c1, c2, w, col, work, th, p, p == 3, fused[:, p:], 6)
    assert float((fused - plain).abs().max() / plain.abs().max()) < 1e-6


def test_loss_and_gradients_match_oracle(me, dc):
    pb = dc["pb"]
    x = torch.as_tensor(dc["x"].astype(np.complex64), device=pb.device).requires_grad_(True)
    th = torch.as_tensor(dc["t"], device=pb.device).requires_grad_(True)
    loss = me.data_consistency_loss(x, th, pb)
    loss.backward()
    ref = O.dc_loss(dc["x"], dc["t"], dc["ksp"], dc["mps"], dc["A"], dc["okg"], dc["orkg"])
    gx, gt = O.dc_loss_grad(dc["x"], dc["t"], dc["ksp"], dc["mps"], dc["A"], dc["okg"], dc["orkg"])
    assert abs(float(loss.detach()) - ref) / ref < 1e-4
    assert _rel(x.grad, gx) < 1e-4
    assert _rel(th.grad, gt) < 1e-4
    assert th.grad.dtype == torch.float64 and x.grad.dtype == torch.complex64


def test_gradients_central_difference_64_directions(me, dc):
    pb = dc["pb"]
    x0 = torch.as_tensor(dc["x"].astype(np.complex64), device=pb.device)
    t0 = torch.as_tensor(dc["t"], dtype=torch.float32, device=pb.device)
    x = x0.clone().requires_grad_(True)
    th = t0.clone().requires_grad_(True)
    me.data_consistency_loss(x, th, pb).backward()
    assert th.grad.dtype == torch.float32
    gen = torch.Generator(device=pb.device)
    gen.manual_seed(43)
    h = 1e-3
    gnorm_x = float(torch.linalg.vector_norm(x.grad))
    gnorm_t = float(torch.linalg.vector_norm(th.grad))
    rel, scaled = [], []
    for _ in range(64):
        dx = torch.randn(SHAPE, dtype=torch.complex64, device=pb.device, generator=gen)
        dt = torch.randn(t0.shape, dtype=torch.float32, device=pb.device, generator=gen)
        with torch.no_grad():
            lp = float(me.data_consistency_loss(x0 + h * dx, t0 + h * dt, pb))
            lm = float(me.data_consistency_loss(x0 - h * dx, t0 - h * dt, pb))
        fd = (lp - lm) / (2 * h)
        an = float(torch.vdot(x.grad.flatten(), dx.flatten()).real) + float((th.grad * dt).sum())
        rel.append(abs(fd - an) / abs(an))
        # error against the scale of the directional derivative's two parts
        # (not blown up by directions nearly orthogonal to the gradient)
        scale = gnorm_x * float(torch.linalg.vector_norm(dx)) + gnorm_t * float(torch.linalg.vector_norm(dt))
        scaled.append(abs(fd - an) / scale)
    print(f"FD vs analytic: median rel {np.median(rel):.2e}, max rel {max(rel):.2e}, max scaled {max(scaled):.2e}")
    assert np.median(rel) < 1e-3
    assert max(scaled) < 1e-3


def test_recompute_off_gives_identical_gradients(me, dc):
    pb = dc["pb"]
    grads = []
    for rec in (True, False):
        x = torch.as_tensor(dc["x"].astype(np.complex64), device=pb.device).requires_grad_(True)
        th = torch.as_tensor(dc["t"], device=pb.device).requires_grad_(True)
        me.data_consistency_loss(x, th, pb, recompute=rec).backward()
        grads.append((x.grad.clone(), th.grad.clone()))
    assert torch.equal(grads[0][0], grads[1][0])
    assert torch.equal(grads[0][1], grads[1][1])


def test_gradient_of_image_only(me, dc):
    pb = dc["pb"]
    x = torch.as_tensor(dc["x"].astype(np.complex64), device=pb.device).requires_grad_(True)
    th = torch.as_tensor(dc["t"], device=pb.device)
    (2.0 * me.data_consistency_loss(x, th, pb)).backward()
    gref = O.dc_loss_grad(dc["x"], dc["t"], dc["ksp"], dc["mps"], dc["A"], dc["okg"], dc["orkg"])[0]
    assert _rel(x.grad, 2.0 * gref) < 1e-4
