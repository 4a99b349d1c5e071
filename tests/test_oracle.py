"""Pins for the CPU oracle (oracle/motion_oracle.py).

The reference ships no tests, fixtures or golden vectors, and cannot run here
(CuPy/SigPy absent; its CG/LM path is CUDA-only).  The oracle is pinned
instead by closed forms, known-answer tests, invariants and finite
differences, each chosen so that a plausible restatement mistake (a wrong
sign, a swapped axis, a dropped term, a reversed order) fails one of them.
"""
import math
import random

import numpy as np
import pytest

from espirit_data import espirit_data
from oracle import motion_oracle as O


def _rand_vol(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex128)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# --------------------------------------------------------------------------- a1
@pytest.mark.parametrize("n,m", [(8, 8), (12, 12), (64, 16), (256, 64), (192, 192)])
def test_grid_closed_form(n, m):
    kg, rkg = O.compute_transform_grids_voxel((n, n, 6), (1.0, 1.0, 1.0), (m, m, 6))
    k = kg["x"].ravel()
    j = np.arange(m)
    kk = 2 * np.pi * (j - m * (j >= m // 2)) / m  # 2*pi*fftfreq with Nyquist at -pi
    assert np.allclose(k, kk, atol=1e-14, rtol=0)
    assert k[m // 2] == pytest.approx(-np.pi)
    r = np.fft.fftshift(np.linspace(-n // 2, n // 2, m, endpoint=False))
    rr = (n / m) * (j - m * (j >= m // 2))
    assert np.allclose(r, rr, atol=1e-12)
    # rk planes are outer products r_b (x) k_a with the documented axis pairs
    _, ry, _ = [np.fft.fftshift(np.linspace(-s // 2, s // 2, d, endpoint=False)) for s, d in ((n, m), (n, m), (6, 6))]
    assert np.allclose(rkg["tan"]["z"], ry.reshape(1, -1, 1) * kg["x"])


# ------------------------------------------------------------------------ a2-a4
@pytest.mark.parametrize("q", [(1, 0, 0), (0, -2, 0), (0, 0, 3), (2, -1, 1)])
def test_integer_translation_is_roll(q):
    shape = (8, 10, 6)
    x = _rand_vol(shape, 1)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    y = O.rigid_apply(x, [*q, 0, 0, 0], kg, rkg)
    ref = np.roll(x, shift=q, axis=(0, 1, 2))
    assert _rel(y, ref) < 1e-13


def _rot90_perm(x, axis):
    """Exact 90-degree rotation about ``axis`` in the fftshift (index 0 =
    origin) frame: point (a, b) -> (-b, a) in the plane (u, v) ordered
    (x1, x2) for axis 0 rotations, (x2, x0) for axis 1, (x0, x1) for axis 2."""
    n0, n1, n2 = x.shape
    if axis == 2:  # z rotation: out[u, v, :] = in[v, -u, :]
        u, v = np.meshgrid(np.arange(n0), np.arange(n1), indexing="ij")
        return x[v % n0, (-u) % n1, :]
    if axis == 0:  # x rotation: out[:, u, v] = in[:, v, -u]
        u, v = np.meshgrid(np.arange(n1), np.arange(n2), indexing="ij")
        return x[:, v % n1, (-u) % n2]
    # y rotation: point (x, z) -> (z, -x): out[u, :, v] = in[-v, :, u]
    u, v = np.meshgrid(np.arange(n0), np.arange(n2), indexing="ij")
    return np.transpose(x[(-v) % n0, :, u % n2], (0, 2, 1))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_rotation_90_is_permutation(axis):
    shape = (8, 8, 8)
    x = _rand_vol(shape, 2)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    params = [0, 0, 0, 0, 0, 0]
    params[3 + axis] = np.pi / 2
    y = O.rigid_apply(x, params, kg, rkg)
    assert _rel(y, _rot90_perm(x, axis)) < 1e-12


def test_unitary_and_adjoint():
    shape = (12, 10, 8)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    x = _rand_vol(shape, 3)
    y = _rand_vol(shape, 4)
    p = [0.7, -1.3, 0.4, 0.11, -0.07, 0.05]
    Tx = O.rigid_apply(x, p, kg, rkg)
    THy = O.rigid_adjoint(y, p, kg, rkg)
    assert abs(np.linalg.norm(Tx) / np.linalg.norm(x) - 1) < 1e-13
    lhs = np.vdot(Tx, y)
    rhs = np.vdot(x, THy)
    assert abs(lhs - rhs) / abs(lhs) < 1e-12
    assert _rel(O.rigid_adjoint(Tx, p, kg, rkg), x) < 1e-12


def _gauss(shape, c, sigma):
    axes = [np.fft.fftfreq(n) * n for n in shape]  # signed coordinate of each index
    g = np.ones(shape)
    for a, (ax, ca) in enumerate(zip(axes, c)):
        d = ax - ca
        d = (d + shape[a] / 2) % shape[a] - shape[a] / 2  # periodic distance
        sh = [1, 1, 1]
        sh[a] = -1
        g = g * np.exp(-0.5 * (d / sigma) ** 2).reshape(sh)
    return g.astype(np.complex128)


def _rot(axis, t):
    c, s = math.cos(t), math.sin(t)
    if axis == 0:  # (y, z) -> (y c - z s, y s + z c)
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    if axis == 1:  # (x, z) -> (x c + z s, -x s + z c)
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


@pytest.mark.parametrize("params", [
    [0, 0, 0, 0.0, 0.0, math.radians(7)],
    [0, 0, 0, math.radians(-6), 0.0, 0.0],
    [0, 0, 0, 0.0, math.radians(5), 0.0],
    [1.5, -2.25, 0.5, math.radians(20), math.radians(-25), math.radians(30)],
])
def test_gaussian_rotation_matches_analytic(params):
    """Off-centre Gaussian (sigma=3, effectively band-limited) versus the
    analytic rotated/translated Gaussian.  With two or more non-zero angles
    this also fixes the composition order Tr Rz Ry Rx (x first)."""
    shape = (48, 48, 48)
    c = np.array([5.0, -4.0, 3.0])
    x = _gauss(shape, c, 3.0)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    y = O.rigid_apply(x, params, kg, rkg)
    R = _rot(2, params[5]) @ _rot(1, params[4]) @ _rot(0, params[3])
    c2 = np.array(params[:3]) + R @ c
    ref = _gauss(shape, c2, 3.0)
    assert _rel(y, ref) < 2e-5
    if sum(1 for t in params[3:] if t) >= 2:  # the reverse order is far off
        R_rev = _rot(0, params[3]) @ _rot(1, params[4]) @ _rot(2, params[5])
        ref_rev = _gauss(shape, np.array(params[:3]) + R_rev @ c, 3.0)
        assert _rel(y, ref_rev) > 0.1


def test_inverse_flag_only_conjugates_translation():
    shape = (8, 8, 8)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    x = _rand_vol(shape, 5)
    p = [1, 2, -1, 0.0, 0.0, 0.0]
    y = O.rigid_apply(x, p, kg, rkg, inverse=True)
    assert _rel(y, np.roll(x, (-1, -2, 1), axis=(0, 1, 2))) < 1e-13


def test_c64_mirror_close_to_f64():
    shape = (16, 16, 12)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    x = _rand_vol(shape, 6)
    p = [0.3, -0.6, 1.1, 0.05, -0.03, 0.04]
    a = O.rigid_apply(x, p, kg, rkg, prec="f64")
    b = O.rigid_apply(x, p, kg, rkg, prec="c64")
    assert b.dtype == np.complex64
    assert 1e-9 < _rel(b, a) < 2e-6


# --------------------------------------------------------------------------- a5
@pytest.mark.parametrize("p_idx", range(6))
def test_derivative_central_difference(p_idx):
    shape = (12, 10, 8)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    x = _gauss(shape, (1.0, -1.0, 0.5), 2.0) * np.exp(1j * 0.3)
    th = np.array([0.4, -0.3, 0.2, 0.12, -0.09, 0.07])
    d = O.rigid_deriv(x, th, kg, rkg, p_idx)
    h = 1e-4
    e = np.zeros(6)
    e[p_idx] = h
    fd = (O.rigid_apply_reverse(x, th + e, kg, rkg) - O.rigid_apply_reverse(x, th - e, kg, rkg)) / (2 * h)
    assert _rel(d, fd) < 1e-6
    # it is NOT the derivative of the forward (x, y, z) order away from 0 (Q1)
    if p_idx >= 3:
        fdf = (O.rigid_apply(x, th + e, kg, rkg) - O.rigid_apply(x, th - e, kg, rkg)) / (2 * h)
        assert _rel(d, fdf) > 1e-3


@pytest.mark.parametrize("p_idx", range(6))
def test_apply_order_derivative_central_difference(p_idx):
    """rigid_deriv_apply is the Jacobian of rigid_apply (x, y, z order) and,
    away from theta = 0, differs from the reference's reverse-order column."""
    shape = (12, 10, 8)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    x = _gauss(shape, (1.0, -1.0, 0.5), 2.0) * np.exp(1j * 0.3) + 0.05 * _rand_vol(shape, 3)
    th = np.array([0.4, -0.3, 0.2, 0.12, -0.09, 0.07])
    d = O.rigid_deriv_apply(x, th, kg, rkg, p_idx)
    h = 1e-4
    e = np.zeros(6)
    e[p_idx] = h
    fd = (O.rigid_apply(x, th + e, kg, rkg) - O.rigid_apply(x, th - e, kg, rkg)) / (2 * h)
    assert _rel(d, fd) < 1e-6
    if p_idx >= 3:
        assert _rel(d, O.rigid_deriv(x, th, kg, rkg, p_idx)) > 1e-3


def test_dc_loss_gradient_central_difference():
    """dc_loss_grad against central differences of dc_loss along random image
    and parameter directions (a dropped term, a wrong conjugate or a sign
    error in either gradient fails this)."""
    shape = (8, 8, 4)
    mps, mask, t, ksp = _small_problem(shape, 2, 3, seed=21, params_scale=0.1)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    x = _rand_vol(shape, 22)
    gx, gt = O.dc_loss_grad(x, t, ksp, mps, mask, kg, rkg)
    rng = np.random.default_rng(23)
    h = 1e-5
    for _ in range(4):
        dx = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        dt = rng.standard_normal(t.shape)
        fd = (O.dc_loss(x + h * dx, t + h * dt, ksp, mps, mask, kg, rkg)
              - O.dc_loss(x - h * dx, t - h * dt, ksp, mps, mask, kg, rkg)) / (2 * h)
        an = float(np.vdot(gx, dx).real) + float(np.sum(gt * dt))
        assert abs(fd - an) < 1e-6 * abs(an)
        # each part on its own
        fdx = (O.dc_loss(x + h * dx, t, ksp, mps, mask, kg, rkg) - O.dc_loss(x - h * dx, t, ksp, mps, mask, kg, rkg)) / (2 * h)
        assert abs(fdx - float(np.vdot(gx, dx).real)) < 1e-6 * abs(fdx)


def test_derivative_out_of_range_returns_input():
    shape = (8, 8, 8)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    x = _rand_vol(shape, 7)
    assert np.array_equal(O.rigid_deriv(x, np.zeros(6) + 0.1, kg, rkg, 7), x)


# ------------------------------------------------------------------------ a6-a8
def _small_problem(shape=(8, 8, 4), ncoils=2, nshots=3, seed=0, params_scale=0.05):
    rng = np.random.default_rng(seed)
    mps = (rng.standard_normal((ncoils, *shape)) + 1j * rng.standard_normal((ncoils, *shape)))
    mps /= np.sqrt(np.sum(np.abs(mps) ** 2, axis=0, keepdims=True))
    lab = rng.integers(0, nshots, size=shape[:2])
    mask = np.stack([(lab == s) for s in range(nshots)])[..., None]
    t = rng.standard_normal((nshots, 6)) * params_scale
    ksp = rng.standard_normal((ncoils, *shape)) + 1j * rng.standard_normal((ncoils, *shape))
    return mps, mask, t, ksp


def _explicit(op, shape):
    n = int(np.prod(shape))
    cols = []
    for i in range(n):
        e = np.zeros(n, dtype=np.complex128)
        e[i] = 1
        cols.append(op(e.reshape(shape)).ravel())
    return np.stack(cols, axis=1)


def test_sum_EhE_is_identity_when_T_is_identity():
    shape = (6, 6, 4)
    mps, mask, t, _ = _small_problem(shape, 3, 4, seed=1)
    t[:] = 0
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    M = np.ones(shape, dtype=bool)
    x = _rand_vol(shape, 9)
    y = O.normal_op(x, mps, mask, t, kg, rkg, M)
    assert _rel(y, x) < 1e-12


def test_normal_op_matches_explicit_matrix_and_cg():
    shape = (6, 6, 4)
    mps, mask, t, ksp = _small_problem(shape, 2, 3, seed=2)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    M = np.ones(shape, dtype=bool)
    # E_s built column by column from basis vectors
    Es = [_explicit(lambda v, s=s: O.encode_forward(v, mps, mask[s], t[s], kg, rkg), shape) for s in range(3)]
    A = sum(E.conj().T @ E for E in Es)
    x = _rand_vol(shape, 10)
    assert _rel(O.normal_op(x, mps, mask, t, kg, rkg, M).ravel(), A @ x.ravel()) < 1e-12
    b = sum(E.conj().T @ (np.asarray(mask[s] * ksp).ravel()) for s, E in enumerate(Es))
    assert _rel(O.cg_rhs(ksp, mps, mask, t, kg, rkg, M).ravel(), b) < 1e-12
    # textbook PCG on the explicit matrix, P = 1/(sum|S|^2 + 1e-3)
    P = (1 / (np.sum(np.abs(mps) ** 2, axis=0) + 1e-3)).ravel()
    xk = np.zeros_like(b)
    r = b - A @ xk
    z = P * r
    p = z.copy()
    rs = np.vdot(r, z).real
    for _ in range(5):
        Ap = A @ p
        al = rs / np.vdot(p, Ap).real
        xk = xk + al * p
        r = r - al * Ap
        z = P * r
        rn = np.vdot(r, z).real
        p = z + (rn / rs) * p
        rs = rn
    xo = O.estimate_image_cg(ksp, mps, mask, t, kg, rkg, P.reshape(shape), M, max_iter=5)
    assert _rel(xo.ravel(), xk) < 1e-10


def test_cg_tolerance_break_updates_x_first():
    """image_estimation.py:146-157: x and r are updated before the break."""
    shape = (6, 6, 4)
    mps, mask, t, ksp = _small_problem(shape, 2, 3, seed=3)
    t[:] = 0
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    M = np.ones(shape, dtype=bool)
    P = np.ones(shape)
    hist = []
    x = O.estimate_image_cg(ksp, mps, mask, t, kg, rkg, P, M, max_iter=10, tol=1e-3, history=hist)
    # A = I here (T = I, partition masks, normalised maps) so CG converges in one step
    assert len(hist) == 1
    b = O.cg_rhs(ksp, mps, mask, t, kg, rkg, M)
    assert _rel(x, b) < 1e-12


# ----------------------------------------------------------------------- a9-a10
def test_lm_jacobian_matches_residual_finite_difference_at_zero():
    """At theta = 0 both rotation orders have the same derivative, so the LM
    Jacobian columns equal finite differences of the residual there."""
    shape = (8, 8, 6)
    mps, mask, _, _ = _small_problem(shape, 2, 2, seed=4)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    x = _gauss(shape, (1, 0, -1), 1.5) + 0.1 * _rand_vol(shape, 11)
    ksp = np.zeros((2, *shape), dtype=np.complex128)
    th = np.zeros(6)
    cols = O.shot_jacobian(mps, mask[0], x, th, kg, rkg)
    h = 1e-5
    for p in range(6):
        e = np.zeros(6)
        e[p] = h
        fd = (O.shot_residual(ksp, mps, mask[0], x, th + e, kg, rkg)
              - O.shot_residual(ksp, mps, mask[0], x, th - e, kg, rkg)) / (2 * h)
        assert _rel(cols[p], fd) < 1e-6


def test_lm_normal_equations_and_solve():
    shape = (8, 8, 4)
    mps, mask, t, ksp = _small_problem(shape, 2, 2, seed=5)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    x = _rand_vol(shape, 12)
    res = O.shot_residual(ksp, mps, mask[0], x, t[0], kg, rkg)
    cols = O.shot_jacobian(mps, mask[0], x, t[0], kg, rkg)
    e0, g, H = O.lm_normal_equations(res, cols, 0.5)
    J = np.stack([c.ravel() for c in cols], axis=1)
    assert e0 == pytest.approx(np.linalg.norm(res) ** 2, rel=1e-12)
    assert np.allclose(g, (J.conj().T @ res.ravel()).real, rtol=1e-12, atol=0)
    assert np.allclose(H, (J.conj().T @ J).real + 0.5 * np.eye(6), rtol=1e-12, atol=1e-12)


def test_lm_translation_only_converges_to_truth():
    """Noise-free, translation-only data: repeated LM steps with the true
    image recover the per-shot translations (zero-mean truth)."""
    shape = (16, 16, 8)
    rng = np.random.default_rng(6)
    nshots = 4
    mps, mask, _, _ = _small_problem(shape, 3, nshots, seed=6)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    x = _gauss(shape, (1, -2, 0), 2.5) + 0.5 * _gauss(shape, (-3, 2, 1), 1.5)
    truth = np.zeros((nshots, 6))
    truth[:, :3] = rng.uniform(-0.8, 0.8, (nshots, 3))
    truth[:, :3] -= truth[:, :3].mean(0)
    ksp = sum(O.encode_forward(x, mps, mask[s], truth[s], kg, rkg) for s in range(nshots))
    t = np.zeros((nshots, 6))
    damp = np.ones(nshots)
    flags = np.zeros(nshots, dtype=bool)
    for _ in range(12):
        O.estimate_transform(ksp, mps, mask, x, kg, rkg, damp, flags, t)
    assert np.max(np.abs(t - truth)) < 1e-4


def test_ga_direction_translation_closed_form():
    """For a translation-only direction v the columns' directional derivative
    is E F^H(-(k.v)^2 F T x): the forward difference of the reference
    (eps = 1e-4) must match it to O(eps)."""
    shape = (8, 8, 6)
    mps, mask, _, _ = _small_problem(shape, 2, 2, seed=31)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    x = _gauss(shape, (0.5, -1, 0.3), 1.5) + 0.05 * _rand_vol(shape, 32)
    th = np.array([0.3, -0.2, 0.1, 0.05, -0.04, 0.03])
    v = np.array([0.7, -0.4, 0.5, 0.0, 0.0, 0.0])
    cols = O.shot_jacobian(mps, mask[0], x, th, kg, rkg)
    jpv = O.ga_direction(mps, mask[0], x, th, cols, v, kg, rkg)
    kv = v[0] * kg["x"] + v[1] * kg["y"] + v[2] * kg["z"]
    # translation derivatives act after the rotations of T_d = Tr Rx Ry Rz
    td = O.rigid_apply_reverse(x, th, kg, rkg)
    dd = np.fft.ifftn(-(kv ** 2) * np.fft.fftn(td, norm="ortho"), norm="ortho")
    ref = np.fft.fftn(mps * dd, axes=(-3, -2, -1), norm="ortho") * mask[0]
    assert _rel(jpv, ref) < 1e-3


def test_ga_step_gate_and_half_step():
    rng = np.random.default_rng(33)
    cols = [rng.standard_normal(20) + 1j * rng.standard_normal(20) for _ in range(6)]
    H = np.array([[np.vdot(a, b).real for b in cols] for a in cols]) + 0.5 * np.eye(6)
    v = rng.standard_normal(6)
    jpv = 1e-3 * (rng.standard_normal(20) + 1j * rng.standard_normal(20))
    rhs = np.array([-np.vdot(jpv, c).real for c in cols])
    a = np.linalg.solve(H, rhs)
    assert np.linalg.norm(a) <= np.linalg.norm(v)
    np.testing.assert_allclose(O.ga_step(H, cols, jpv, v), v + 0.5 * a, rtol=1e-12)
    big = 1e6 * jpv  # acceleration larger than the step: rejected
    np.testing.assert_array_equal(O.ga_step(H, cols, big, v), v)


def test_lm_damping_rules():
    """transform_estimation.py:138-144: accept -> damp/5 (floor 1e-4),
    reject -> damp*1.5 (cap 1e16)."""
    shape = (8, 8, 4)
    mps, mask, _, _ = _small_problem(shape, 2, 1, seed=7)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    x = _gauss(shape, (0, 1, 0), 2.0)
    truth = np.array([[0.3, -0.2, 0.1, 0, 0, 0]])
    ksp = O.encode_forward(x, mps, mask[0], truth[0], kg, rkg)
    t = np.zeros((1, 6))
    damp = np.array([1e-4 * 3])
    flags = np.zeros(1, dtype=bool)
    trace = []
    O.estimate_transform(ksp, mps, mask, x, kg, rkg, damp, flags, t, trace=trace)
    assert trace[0][2] and damp[0] == pytest.approx(1e-4)
    # already at the optimum: any step is rejected (e1 >= e0) and damp grows
    t2 = np.zeros((1, 6))
    damp2 = np.array([1e16])
    ksp0 = O.encode_forward(x, mps, mask[0], t2[0], kg, rkg)
    trace2 = []
    O.estimate_transform(ksp0, mps, mask, x, kg, rkg, damp2, flags, t2, trace=trace2)
    assert not trace2[0][2] and damp2[0] == 1e16


# -------------------------------------------------------------------------- a15
def _tile_invariants(order, K, tile):
    order = np.asarray(order)
    lines, parts = order[:, 0], order[:, 1]
    assert len(set(zip(lines.tolist(), parts.tolist()))) == len(order)  # no duplicates
    assert lines.max() < K[1] and parts.max() < K[0]


@pytest.mark.parametrize("K,tile,accel", [((64, 64), (4, 4), 1), ((128, 128), (8, 8), 2), ((256, 256), (16, 16), 2), ((48, 64), (4, 8), 2)])
def test_disorder_invariants(K, tile, accel):
    cfg = O.DisorderConfig(max_line_number=K[1] - 1, max_partition_number=K[0] - 1,
                           pat_lines_to_measure=K[1] // accel, block_lin=tile[1], block_par=tile[0])
    order, (L, P) = O.make_disorder_order(cfg)
    assert (L, P) == (K[1], K[0])
    assert len(order) == K[0] * K[1] // accel
    _tile_invariants(order, K, tile)
    # every tile receives the same number of samples
    o = np.asarray(order)
    tiles = (o[:, 1] // tile[0]) * (K[1] // tile[1]) + (o[:, 0] // tile[1])
    cnt = np.bincount(tiles)
    assert cnt.min() == cnt.max()
    nshots = tile[0] * tile[1] // accel
    A = O.disorder_masks(order, nshots, K)
    assert A.shape == (nshots, *K, 1)
    assert int(A.sum()) == len(order) - len(order) % nshots
    assert A.sum(0).max() == 1  # shots are disjoint


def test_disorder_first_draws_follow_stdlib_shuffle():
    """The first tile's draws are exactly ``random.Random(seed).shuffle`` of
    range(block_lin*block_par) mapped through the interleave formula."""
    cfg = O.DisorderConfig(15, 15, 16, 4, 4, seed=123456)
    order, _ = O.make_disorder_order(cfg)
    rng = random.Random(123456)
    idx = list(range(16))
    rng.shuffle(idx)
    n_groups = 16
    for j in range(16):
        loc = idx[j]
        assert order[j * n_groups] == (loc % 4, loc // 4)


# -------------------------------------------------------------------------- a13
def test_sp_resize_and_centered_fft():
    a = np.arange(8 * 6).reshape(8, 6)
    b = O.sp_resize(a, (4, 2))
    assert np.array_equal(b, a[2:6, 2:4])
    c = O.sp_resize(b, (8, 6))
    assert np.array_equal(c[2:6, 2:4], b) and c.sum() == b.sum()
    d = np.zeros((8, 8), dtype=complex)
    d[4, 4] = 1.0  # delta at the centre -> flat spectrum
    f = O.sp_fft(d, (0, 1))
    assert np.allclose(f, 1 / 8)
    assert np.allclose(O.sp_ifft(f, (0, 1)), d)


# ------------------------------------------------------------------ SENSE (f1)
def _sense_problem(shape, ncoils, seed, frac=0.5):
    rng = np.random.default_rng(seed)
    mps = rng.standard_normal((ncoils, *shape)) + 1j * rng.standard_normal((ncoils, *shape))
    mps /= np.sqrt(np.sum(np.abs(mps) ** 2, axis=0, keepdims=True))
    lines = rng.random(shape[:2]) < frac
    y = (rng.standard_normal((ncoils, *shape)) + 1j * rng.standard_normal((ncoils, *shape))) * lines[None, ..., None]
    return mps, lines, y


def test_sense_full_sampling_is_one_step():
    shape = (6, 4, 4)
    mps, _, _ = _sense_problem(shape, 3, 50)
    rng = np.random.default_rng(51)
    y = rng.standard_normal((3, *shape)) + 1j * rng.standard_normal((3, *shape))
    info = {}
    x = O.sense_recon(y, mps, max_iter=5, tol=1e-10, info=info)
    direct = np.sum(np.conj(mps) * O._ifftc(y, (-3, -2, -1)), axis=0)
    assert _rel(x, direct) < 1e-12
    assert info["iterations"] == 1  # r = 0 (to rounding) after the first step


def test_sense_matches_explicit_solve():
    shape = (4, 4, 2)
    mps, lines, y = _sense_problem(shape, 2, 52)
    n = int(np.prod(shape))
    w = np.broadcast_to(lines[..., None], shape).astype(float)
    sw = np.sqrt(w)
    A = np.zeros((2 * n, n), dtype=np.complex128)
    for j in range(n):
        e = np.zeros(n, dtype=np.complex128)
        e[j] = 1
        A[:, j] = (sw * O._fftc(mps * e.reshape(shape), (-3, -2, -1))).ravel()
    lam = 0.3
    AHA = A.conj().T @ A + lam * np.eye(n)
    ref = np.linalg.solve(AHA, A.conj().T @ (sw * y).ravel()).reshape(shape)
    P = 1.0 / (np.sum(np.abs(mps) ** 2, axis=0) + 0.5)
    x = O.sense_recon(y, mps, lamda=lam, P=P, max_iter=4 * n)
    assert _rel(x, ref) < 1e-8


def test_sense_zero_data_stops_before_first_step():
    shape = (4, 4, 2)
    mps, lines, _ = _sense_problem(shape, 2, 53)
    info = {}
    x0 = _rand_vol(shape, 54)
    y = np.zeros((2, *shape), dtype=np.complex128)
    w = np.broadcast_to(lines[..., None], shape).astype(float)
    # b = 0 and x0 = 0: resid = 0 <= tol before any step
    x = O.sense_recon(y, mps, weights=w, max_iter=10, info=info)
    assert info["iterations"] == 0 and np.all(x == 0)
    # a nonzero start does iterate
    O.sense_recon(y, mps, weights=w, max_iter=10, x0=x0, info=info)
    assert info["iterations"] > 0


# ------------------------------------------------------- calibration (f3)
def test_noise_whiten_factor():
    rng = np.random.default_rng(70)
    C = 5
    mix = rng.standard_normal((C, C)) + 1j * rng.standard_normal((C, C))
    noise = mix @ (rng.standard_normal((C, 400)) + 1j * rng.standard_normal((C, 400)))
    ksp = rng.standard_normal((C, 3, 4)) + 1j * rng.standard_normal((C, 3, 4))
    kw, W = O.noise_whiten(ksp, noise, reg=1e-3)
    Lr = np.linalg.inv(W.conj().T)
    assert np.allclose(Lr, np.tril(Lr), atol=1e-12)
    assert np.all(np.abs(np.diag(Lr).imag) < 1e-12) and np.all(np.diag(Lr).real > 0)
    cov = noise @ noise.conj().T / noise.shape[1] + 1e-3 * np.eye(C)
    assert np.allclose(Lr @ Lr.conj().T, cov, rtol=1e-10, atol=1e-10)
    assert np.allclose(kw, np.einsum("ij,jab->iab", W, ksp))


def test_coil_compression_index_rule_and_subspace():
    rng = np.random.default_rng(71)
    shape = (6, 5, 4)
    b = rng.standard_normal((2, 4)) + 1j * rng.standard_normal((2, 4))  # 4 coils, rank 2
    f = rng.standard_normal((2, *shape)) + 1j * rng.standard_normal((2, *shape))
    mps = np.einsum("kc,kxyz->cxyz", b, f)
    cc, _, A, S = O.coil_compression(mps, perc=0.999)
    share = np.cumsum(S) / np.sum(S)
    assert S[2] < 1e-9 * S[0]  # rank 2
    assert share[0] < 0.999 <= share[1]
    assert A.shape == (1, 4)  # argmax index 1: one coil kept (the reference's rule)
    # A's row is the dominant eigenvector of P
    m = mps.reshape(4, -1)
    norm = 1 / (np.sum(np.abs(mps) ** 2, axis=0).ravel() + 1e-6)
    P = (m * norm) @ m.conj().T
    v = A[0].conj()
    assert np.allclose(P @ v, S[0] * v, rtol=1e-9, atol=1e-9 * S[0])
    assert np.allclose(cc[0], np.einsum("c,cxyz->xyz", A[0], mps))


def test_hankel_structure_single_sample():
    calib = np.zeros((7, 6, 5, 3), dtype=np.complex128)
    calib[4, 2, 3, 1] = 2 - 1j
    H = O.fold_rx(O.hankel_fwd(calib, (3, 3, 3)))
    K = 27
    nbx, nby = 5, 4
    rows, cols = np.nonzero(H)
    assert len(rows) > 0
    for r, b in zip(rows, cols):
        k, c = r % K, r // K
        ix, iy, iz = k % 3, (k // 3) % 3, k // 9  # Fortran order inside the block
        bx, by, bz = b % nbx, (b // nbx) % nby, b // (nbx * nby)  # kx_start fastest
        assert c == 1 and (bx + ix, by + iy, bz + iz) == (4, 2, 3)
        assert H[r, b] == 2 - 1j
    # every block that covers the sample has exactly one entry
    covering = sum(1 for bz in range(3) for by in range(4) for bx in range(5)
                   if 0 <= 4 - bx < 3 and 0 <= 2 - by < 3 and 0 <= 3 - bz < 3)
    assert len(rows) == covering


def test_espirit_recovers_known_maps():
    imsize = (24, 20, 16)
    calib, maps, obj = espirit_data(imsize)
    info = {}
    sens = O.rx_espirit_3d(calib, imsize, kernel_size=(3, 3, 3), info=info)
    assert 0 < info["kept"] < calib.shape[3] * 27
    inner = obj > 0.3 * obj.max()
    s = sens[inner]  # (n, nc)
    t = np.transpose(maps, (1, 2, 3, 0))[inner]
    corr = np.abs(np.sum(np.conj(s) * t, axis=1)) / (np.linalg.norm(s, axis=1) * np.linalg.norm(t, axis=1))
    assert np.median(corr) > 0.999 and corr.min() > 0.98
    assert np.allclose(np.linalg.norm(s, axis=1), 1, atol=1e-9)
    assert np.all(np.abs(s[:, 0].imag) < 1e-9) and np.all(s[:, 0].real >= 0)


def test_gradient_entropy_constant_volume():
    """All gradients zero: every voxel in bin 0 of [0, 1e-8]; density 1/db0."""
    v = np.full((5, 4, 3), 2.5, dtype=np.float32)
    ent, counts = O.gradient_entropy(v)
    assert counts[0] == v.size and counts[1:].sum() == 0
    hi = np.float32(0) + 1e-8
    edges = np.linspace(0, hi, 257, dtype=np.result_type(0, hi, np.float32))
    db = np.diff(edges).astype(np.float64)
    h = np.full(256, 1e-12)
    h[0] += 1.0 / db[0]
    assert np.isclose(ent, -np.sum(h * np.log2(h)), rtol=1e-12)


def test_gradient_entropy_ramp():
    """A ramp of slope 0.5 along axis 1: gradient 0.5 except the last y-plane
    (0); 0.5 + 1e-8 rounds to 0.5 in float32, so 0.5 falls on the last edge
    and goes to the last bin."""
    n0, n1, n2 = 3, 6, 4
    v = np.broadcast_to((0.5 * np.arange(n1, dtype=np.float32))[None, :, None], (n0, n1, n2)).copy()
    ent, counts = O.gradient_entropy(v)
    zero = n0 * n2
    assert counts[0] == zero and counts[255] == v.size - zero and counts[1:255].sum() == 0
    db = 0.5 / 256
    p0, p1 = zero / v.size / db, (v.size - zero) / v.size / db
    h = np.full(256, 1e-12)
    h[0] += p0
    h[255] += p1
    assert np.isclose(ent, -np.sum(h * np.log2(h)), rtol=1e-6)


def test_error_volume_known():
    a = np.array([3 + 4j, 1 + 0j, 0 + 2j])
    b = np.array([0 + 0j, 1 + 0j, 0 - 1j])
    e, n = O.error_volume(a, b)
    assert np.allclose(e, [5, 0, 3]) and np.isclose(n, np.sqrt(34))


def test_shot_masks_from_counters_known():
    lines = [0, 1, 2, 3, 4, 5, 6]
    parts = [1, 1, 0, 0, 2, 2, 3]
    m = O.shot_masks_from_counters(lines, parts, 3, 4, 8)
    assert m.shape == (3, 4, 8, 1)
    assert sorted(map(tuple, np.argwhere(m[..., 0]).tolist())) == [(0, 1, 0), (0, 1, 1), (1, 0, 2), (1, 0, 3),
                                                                    (2, 2, 4), (2, 2, 5)]  # sample 6 dropped
