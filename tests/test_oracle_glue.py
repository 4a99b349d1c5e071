"""Pins for the oracle's composition glue: generate_corrupted_kspace
(simulation.py:65-85), JointRecon (joint_recon.py:38-115) and
pyramid_reconstruction (pyramid.py:24-118).

Each answer here comes from a closed form or a brute-force sum that does not
reuse the oracle's own shift/FFT calls:

* the centred orthonormal DFT is written as an explicit matrix,
  D[j, i] = exp(-2 pi i (j - n/2)(i - n/2) / n) / sqrt(n), so the
  ifftshift/fftshift conventions of the glue are checked independently;
* with zero motion, masks that partition k-space and sum_c |S_c|^2 = 1 the
  normal operator is the identity, so one CG step returns the image exactly;
* one coil with S = 1 and half of k-space sampled makes the operator a
  projection, so one CG step returns the zero-filled reconstruction;
* data outside the range of every shot's E_s change neither the image nor
  the motion, and their energy is the objective;
* with a piecewise-constant sum_c |S_c|^2 = w the operator is diag(M w) and
  PCG converges in as many steps as PA has distinct values;
* the pyramid's per-level JointRecon is replaced by a recorder, which fixes
  the level shapes, grids, iteration counts, the unscaled carry of t and the
  motion the final CG level uses.
"""
import numpy as np
import pytest

from oracle import motion_oracle as O


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _cdft(n):
    j = np.arange(n) - n // 2
    return np.exp(-2j * np.pi * np.outer(j, j) / n) / np.sqrt(n)


def _cfft(x, axes, inverse=False):
    """Brute-force centred orthonormal DFT (index n//2 is the origin)."""
    x = np.asarray(x, dtype=np.complex128)
    for ax in axes:
        D = _cdft(x.shape[ax])
        if inverse:
            D = D.conj().T
        x = np.moveaxis(np.tensordot(D, np.moveaxis(x, ax, 0), axes=(1, 0)), 0, ax)
    return x


def _vol(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _unit_maps(shape, ncoils, seed):
    """Random smooth-free maps with sum_c |S_c|^2 = 1 at every voxel."""
    m = _vol((ncoils, *shape), seed)
    return m / np.sqrt(np.sum(np.abs(m) ** 2, axis=0))


def _partition_masks(shape2, nshots, seed, keep=1.0):
    """(S, N0, N1, 1) bool: every (k0, k1) of a random ``keep`` fraction
    belongs to exactly one shot."""
    rng = np.random.default_rng(seed)
    owner = rng.integers(0, nshots, size=shape2)
    sampled = rng.random(shape2) < keep
    A = np.stack([(owner == s) & sampled for s in range(nshots)])
    return A[..., None]


def test_brute_force_dft_is_numpys_centred_fft():
    x = _vol((8, 6, 4), 0)
    ref = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x), norm="ortho"))
    assert _rel(_cfft(x, (0, 1, 2)), ref) < 1e-13


# ------------------------------------------------------------- simulation.py:65-85
def test_corrupted_kspace_zero_motion_is_masked_centred_dft():
    shape, C, S = (8, 6, 4), 2, 3
    gt = _vol(shape, 1)
    mps = _vol((C, *shape), 2)
    A = _partition_masks(shape[:2], S, 3, keep=0.5)
    ksp = O.generate_corrupted_kspace(gt, mps, A, np.zeros((S, 6)))
    expected = A.any(0)[None] * _cfft(mps * gt[None], (1, 2, 3))
    assert _rel(ksp, expected) < 1e-12


def test_corrupted_kspace_takes_each_shots_motion_on_its_own_samples():
    shape, C = (8, 6, 4), 2
    gt = _vol(shape, 4)
    mps = _vol((C, *shape), 5)
    A = _partition_masks(shape[:2], 2, 6)
    q = [(1, 0, -1), (0, 2, 1)]
    t = np.array([[*q[0], 0, 0, 0], [*q[1], 0, 0, 0]], dtype=float)
    ksp = O.generate_corrupted_kspace(gt, mps, A, t)
    # an integer translation is a circular shift of the centred volume as well
    expected = sum(A[s][None] * _cfft(mps * np.roll(gt, q[s], axis=(0, 1, 2))[None], (1, 2, 3)) for s in range(2))
    assert _rel(ksp, expected) < 1e-12


# ------------------------------------------------------------- joint_recon.py:38-115
def _data(shape, C, S, seed, maps=None, motion=None, keep=1.0):
    gt = _vol(shape, seed)
    mps = _unit_maps(shape, C, seed + 1) if maps is None else maps
    A = _partition_masks(shape[:2], S, seed + 2, keep)
    t = np.zeros((S, 6)) if motion is None else motion
    return gt, mps, A, O.generate_corrupted_kspace(gt, mps, A, t)


def test_joint_recon_identity_operator_one_cg_step():
    """Zero motion, masks partitioning k-space, sum|S|^2 = 1: A = I, so one CG
    step returns the image; the LM step leaves t at 0; the objective of the
    exact image is 0; max_norm = max |S_c g|; the output is rescaled and
    recentred to the ground truth."""
    shape, C, S = (8, 8, 4), 3, 4
    gt, mps, A, ksp = _data(shape, C, S, 10)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    app = O.JointRecon(ksp, mps, A, kg, rkg, max_cg_iter=1, max_joint_iter=1)
    assert app.max_norm == pytest.approx(float(np.abs(mps * gt[None]).max()), rel=1e-12)
    assert app.M.all() and np.allclose(app.P, 1 / 1.001, rtol=1e-12)
    x, t = app.run()
    assert _rel(x, gt) < 1e-10
    assert np.abs(t).max() < 1e-10
    y2 = float(np.sum(np.abs(ksp) ** 2)) / app.max_norm ** 2
    assert app.objective_history.shape == (1,) and app.objective_history[0] < 1e-20 * y2


def test_joint_recon_zero_filled_projection():
    """One coil with S = 1, two shots over half of k-space, zero motion:
    A = F^H M F is a projection, so one CG step returns the zero-filled
    reconstruction F^H M F g (brute-force centred DFT); the shifts of ksp and
    of the masks (joint_recon.py:63-65) must agree for this to hold."""
    shape, S = (8, 8, 4), 2
    gt, mps, A, ksp = _data(shape, 1, S, 50, maps=np.ones((1, 8, 8, 4), complex), keep=0.5)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    app = O.JointRecon(ksp, mps, A, kg, rkg, max_cg_iter=1, max_joint_iter=1)
    x, t = app.run()
    expected = _cfft(A.any(0) * _cfft(gt, (0, 1, 2)), (0, 1, 2), inverse=True)
    assert _rel(x, expected) < 1e-10
    assert np.abs(t).max() < 1e-10


def _explicit_shot_operator(mps, mask_s):
    """E_s (zero motion, centred frame) as a matrix: v -> mask_s * D(S_c v)."""
    C, *shape = mps.shape
    N = int(np.prod(shape))
    cols = []
    for i in range(N):
        e = np.zeros(N, complex)
        e[i] = 1
        cols.append((mask_s[None] * _cfft(mps * e.reshape(shape)[None], (1, 2, 3))).ravel())
    return np.stack(cols, axis=1)


def test_joint_recon_objective_of_data_outside_the_range():
    """Data y + d with d_s in the null space of E_s^H on shot s's samples: the
    right-hand side, the CG image (g) and the LM gradient (zero) are those of
    y alone, t stays 0, and the objective is ||d||^2 / max_norm^2."""
    shape, C, S = (8, 8, 4), 3, 2
    gt, mps, A, ksp = _data(shape, C, S, 60)
    rng = np.random.default_rng(61)
    d = np.zeros_like(ksp)
    for s in range(S):
        E = _explicit_shot_operator(mps, A[s])
        ds = (rng.standard_normal(ksp.shape) + 1j * rng.standard_normal(ksp.shape)) * A[s][None]
        ds = ds.ravel() - E @ np.linalg.lstsq(E, ds.ravel(), rcond=None)[0]
        assert np.linalg.norm(E.conj().T @ ds) < 1e-10 * np.linalg.norm(ds)
        d += ds.reshape(ksp.shape)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    app = O.JointRecon(ksp + d, mps, A, kg, rkg, max_cg_iter=1, max_joint_iter=1)
    x, t = app.run()
    assert _rel(x, gt) < 1e-10
    assert np.abs(t).max() < 1e-12
    assert app.objective_history[0] == pytest.approx(float(np.sum(np.abs(d) ** 2)) / app.max_norm ** 2, rel=1e-9)


def _piecewise_problem(signal_off_support=True):
    """sum_c |S_c|^2 = w with w = 0.05 on the first two x0 rows, 0.5 on the
    next two and 1 elsewhere (centred frame); optionally no signal on the
    w = 0.05 rows."""
    shape, C, S = (8, 8, 4), 3, 4
    w = np.ones(shape)
    w[:2] = 0.05
    w[2:4] = 0.5
    mps = _unit_maps(shape, C, 21) * np.sqrt(w)[None]
    gt = _vol(shape, 20)
    if not signal_off_support:
        gt[:2] = 0
    A = _partition_masks(shape[:2], S, 22)
    return shape, w, gt, mps, A, O.generate_corrupted_kspace(gt, mps, A, np.zeros((S, 6)))


def test_joint_recon_preconditioner_and_support():
    """P = 1/(w + 1e-3) and M = w > 0.1 in the uncentred frame; A = diag(M w)
    with two distinct values of PA on M, so PCG ends after two steps with the
    image itself (no signal off M here, so LM has nothing to fit)."""
    shape, w, gt, mps, A, ksp = _piecewise_problem(signal_off_support=False)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    app = O.JointRecon(ksp, mps, A, kg, rkg, max_cg_iter=3, max_joint_iter=1)
    wu = np.fft.ifftshift(w)
    assert np.allclose(app.P, 1 / (wu + 1e-3), rtol=1e-12)
    assert np.array_equal(app.M, wu > 0.1)
    x, t = app.run()
    assert _rel(x, gt) < 1e-10
    assert np.abs(t).max() < 1e-10
    assert app.objective_history[0] < 1e-20 * float(np.sum(np.abs(ksp) ** 2)) / app.max_norm ** 2


def test_joint_recon_masks_after_reregistration():
    """A common translation of every shot is moved into the image by the mean
    re-registration (transform_estimation.py:147-150); x *= M
    (joint_recon.py:100) then removes what it shifted out of the support, so
    the result is zero off M and t is zero-mean."""
    shape, w, gt, mps, A, ksp = _piecewise_problem()
    S = A.shape[0]
    t0 = np.tile([1.0, 0, 0, 0, 0, 0], (S, 1))
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    app = O.JointRecon(ksp, mps, A, kg, rkg, t0=t0, max_cg_iter=3, max_joint_iter=1)
    x, t = app.run()
    Mc = w > 0.1
    assert np.abs(x[~Mc]).max() == 0.0
    assert np.abs(x[Mc]).max() > 0.1 * np.abs(gt).max()
    assert np.abs(t.mean(0)).max() < 1e-12
    # without the support mask the re-registered image leaks off M
    shifted = np.roll(gt * Mc, 1, axis=0)
    assert np.abs(shifted[~Mc]).max() > 0


# ------------------------------------------------------------- pyramid.py:24-118
class _LevelRecorder:
    """Stands in for the per-level JointRecon: records what each level is
    given and returns a fixed motion estimate per level."""

    def __init__(self, results):
        self.results = results
        self.calls = []

    def __call__(self, ksp, mps, A, kgrid, rkgrid, t0=None, max_joint_iter=100, **kw):
        rec = self
        rec.calls.append(dict(ksp=ksp, mps=mps, A=A, kgrid=kgrid, rkgrid=rkgrid,
                              t0=None if t0 is None else np.array(t0, copy=True), iters=max_joint_iter, kw=kw))
        idx = len(rec.calls) - 1

        class _Run:
            def run(self_inner):
                return None, rec.results[idx].copy()

        return _Run()


@pytest.mark.parametrize("sampling", ["disorder", "sequential"])
def test_pyramid_levels_carry_and_final_cg(monkeypatch, sampling):
    shape, C, S = (16, 16, 4), 2, 4
    t1 = np.tile([0.3, -0.2, 0.1, 0.01, -0.02, 0.015], (S, 1))  # returned by the x4 level
    t2 = np.tile([1.0, -0.5, 0.25, 0.03, 0.02, -0.01], (S, 1))  # returned by the x2 level
    # data moved by t2 (the same for every shot): the final CG level sees the
    # identity operator T^H T and must return gt exactly when it uses t2
    gt, mps, A, ksp = _data(shape, C, S, 30, motion=t2)
    rec = _LevelRecorder([t1, t2])
    monkeypatch.setattr(O, "JointRecon", rec)
    x, t = O.pyramid_reconstruction(ksp, mps, A, S, n_spatial_levels=3, n_joint_iters=5, sampling=sampling)
    assert len(rec.calls) == 2
    for c, f in zip(rec.calls, (4, 2)):
        ds = (shape[0] // f, shape[1] // f if sampling == "disorder" else shape[1], shape[2])
        assert c["ksp"].shape == (C, *ds) and c["mps"].shape == (C, *ds) and c["A"].shape == (S, *ds[:2], 1)
        assert c["iters"] == 5 * f
        # grids of the level: k in level samples, r in full-resolution voxels
        kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1), ds)
        for a in "xyz":
            assert np.array_equal(c["kgrid"][a], kg[a])
            assert np.array_equal(c["rkgrid"]["tan"][a], rkg["tan"][a])
    assert np.array_equal(rec.calls[0]["t0"], np.zeros((S, 6)))
    assert np.array_equal(rec.calls[1]["t0"], t1)  # carried unscaled
    assert np.array_equal(t, t2)
    assert _rel(x, gt) < 1e-9


def test_pyramid_level_resampling_known_answers():
    """pyramid.py:65-67 with SigPy's centred ortho fft and resize: constant
    maps c come out as c * sqrt(N / n) (the spectrum keeps its DC value while
    the level has n of the N voxels); the k-space DC sample stays at the
    level's centre; the masks are centre-cropped."""
    shape, C, S = (16, 12, 4), 2, 3
    mps = np.full((C, *shape), 0.5 + 0.25j)
    ksp = np.zeros((C, *shape), dtype=complex)
    ksp[:, shape[0] // 2, shape[1] // 2, shape[2] // 2] = [1.0, 2.0j]
    A = _partition_masks(shape[:2], S, 40)
    for sampling in ("disorder", "sequential"):
        for f, k_res, m_res, a_res, ds in O.pyramid_levels(ksp, mps, A, S, 3, sampling):
            n_ratio = np.prod(shape) / np.prod(ds)
            assert np.allclose(m_res, (0.5 + 0.25j) * np.sqrt(n_ratio), atol=1e-12)
            nz = np.argwhere(np.abs(k_res) > 0)
            assert nz.tolist() == [[0, ds[0] // 2, ds[1] // 2, ds[2] // 2], [1, ds[0] // 2, ds[1] // 2, ds[2] // 2]]
            o0, o1 = shape[0] // 2 - ds[0] // 2, shape[1] // 2 - ds[1] // 2
            assert np.array_equal(a_res, A[:, o0:o0 + ds[0], o1:o1 + ds[1]])
