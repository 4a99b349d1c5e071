"""Differentiation at 192 x 192 x 128 with 24 coils and 48 DISORDER shots
(tile 8 x 12, acceleration 2; simulation.make_config("c3")): the loss
L = 0.5 sum_s ||M_s F (S T(theta_s) x) - y_s||^2 and its hand-written
gradients with respect to the image and the 48 x 6 motion.

* Central differences at step 1e-3 along 64 random (image, motion)
  directions: every error below 1e-3 of |g_x||dx| + |g_theta||dtheta| (the
  plain relative error means little for directions nearly orthogonal to the
  gradient) and the median relative error below 1e-2.  At this size a step of
  1e-3 rad moves the edge voxels by about 0.1 voxel, so the difference
  quotient carries an O(h^2) truncation term; Richardson extrapolation from
  steps h and h/2 removes it, and what remains must be below 3e-5 of the same
  scale.
* Image-only directions with a component along the image gradient: L is
  quadratic in x, so the central difference is exact up to rounding;
  relative error 1e-4.
* Rotations zero (translations not): the apply order Tr Rz Ry Rx and the
  reference derivative class's order Tr Rx Ry Rz have the same Jacobian, so
  the motion gradient equals the LM gradient Re <r, J_p> that
  SenseProblem.lm_step forms from the reference-order columns
  (RigidTransformDerivativeCuda, transform.py:476-504): relative 1e-4.
* recompute on and off: identical gradients.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3(me):
    dev = torch.device("cuda:0")
    d = me.simulation.make_config("c3", device=dev)
    cfg = d["config"]
    shape = tuple(cfg["shape"])
    assert shape == (192, 192, 128) and cfg["coils"] == 24 and d["A"].shape[0] == 48
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    pb = me.encoding.SenseProblem(d["ksp"], d["mps"], d["A"], kg, rkg, centered=True)
    x0 = torch.fft.ifftshift(d["gt"].to(dev), dim=(0, 1, 2)).to(torch.complex64)
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)
    t0 = 0.01 * torch.randn((48, 6), dtype=torch.float32, device=dev, generator=gen)
    print(f"c3: {shape}, {cfg['coils']} coils, 48 shots", flush=True)
    return dict(d, pb=pb, x0=x0, t0=t0, gen=gen, shape=shape, kg=kg, rkg=rkg)


def _grads(me, pb, x0, t0, recompute=True):
    x = x0.clone().requires_grad_(True)
    th = t0.clone().requires_grad_(True)
    loss = me.data_consistency_loss(x, th, pb, recompute=recompute)
    loss.backward()
    return float(loss.detach()), x.grad, th.grad


def test_c3_central_differences_64_directions(me, c3):
    pb, x0, t0, gen = c3["pb"], c3["x0"], c3["t0"], c3["gen"]
    _, gx, gt = _grads(me, pb, x0, t0)
    gnx, gnt = float(torch.linalg.vector_norm(gx)), float(torch.linalg.vector_norm(gt))
    h = 1e-3

    def cd(dx, dt, step):
        return (float(me.data_consistency_loss(x0 + step * dx, t0 + step * dt, pb))
                - float(me.data_consistency_loss(x0 - step * dx, t0 - step * dt, pb))) / (2 * step)

    rel, scaled, scaled_rich = [], [], []
    with torch.no_grad():
        for _ in range(64):
            dx = torch.randn(x0.shape, dtype=torch.complex64, device=x0.device, generator=gen)
            dt = torch.randn(t0.shape, dtype=torch.float32, device=x0.device, generator=gen)
            fd, fd2 = cd(dx, dt, h), cd(dx, dt, h / 2)
            an = float(torch.vdot(gx.flatten(), dx.flatten()).real) + float((gt * dt).sum())
            scale = gnx * float(torch.linalg.vector_norm(dx)) + gnt * float(torch.linalg.vector_norm(dt))
            rel.append(abs(fd - an) / abs(an))
            scaled.append(abs(fd - an) / scale)
            # Richardson: (4 D(h/2) - D(h)) / 3 cancels the h^2 truncation term
            scaled_rich.append(abs((4 * fd2 - fd) / 3 - an) / scale)
    print(f"c3 FD: median rel {np.median(rel):.2e}, max rel {max(rel):.2e}, max scaled {max(scaled):.2e}, "
          f"max scaled after Richardson {max(scaled_rich):.2e}", flush=True)
    assert np.median(rel) < 1e-2
    assert max(scaled) < 1e-3
    assert max(scaled_rich) < 3e-5


def test_c3_image_gradient_quadratic_directions(me, c3):
    pb, x0, t0, gen = c3["pb"], c3["x0"], c3["t0"], c3["gen"]
    _, gx, _ = _grads(me, pb, x0, t0)
    xn = float(torch.linalg.vector_norm(x0))
    errs = []
    with torch.no_grad():
        for _ in range(4):
            r = torch.randn(x0.shape, dtype=torch.complex64, device=x0.device, generator=gen)
            dx = gx / torch.linalg.vector_norm(gx) + r / torch.linalg.vector_norm(r)
            dx = dx * (0.05 * xn / float(torch.linalg.vector_norm(dx)))
            fd = (float(me.data_consistency_loss(x0 + dx, t0, pb)) - float(me.data_consistency_loss(x0 - dx, t0, pb))) / 2
            an = float(torch.vdot(gx.flatten(), dx.flatten()).real)
            errs.append(abs(fd - an) / abs(an))
    print(f"c3 image-only FD: max rel {max(errs):.2e}", flush=True)
    assert max(errs) < 1e-4


def test_c3_motion_gradient_equals_reference_order_lm_gradient(me, c3):
    pb, x0 = c3["pb"], c3["x0"]
    gen = torch.Generator(device=x0.device)
    gen.manual_seed(6)
    t0 = torch.zeros((48, 6), dtype=torch.float64, device=x0.device)
    t0[:, :3] = 0.5 * torch.randn((48, 3), dtype=torch.float64, device=x0.device, generator=gen)
    _, _, gt = _grads(me, pb, x0, t0)
    d = c3
    lm = me.encoding.SenseProblem(d["ksp"], d["mps"], d["A"], d["kg"], d["rkg"], centered=True)
    lm.theta.copy_(t0)
    lm.lm_step(lm.image_to_plane(x0))
    g_lm = lm._stats[:, 1:7]
    err = float(torch.linalg.vector_norm(gt - g_lm) / torch.linalg.vector_norm(g_lm))
    print(f"c3 motion gradient vs reference-order LM gradient: rel {err:.2e}", flush=True)
    assert err < 1e-4


def test_c3_recompute_on_off_identical(me, c3):
    pb, x0, t0 = c3["pb"], c3["x0"], c3["t0"]
    l1, gx1, gt1 = _grads(me, pb, x0, t0, recompute=True)
    l2, gx2, gt2 = _grads(me, pb, x0, t0, recompute=False)
    assert l1 == l2
    assert torch.equal(gx1, gx2) and torch.equal(gt1, gt2)
