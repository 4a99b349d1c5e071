"""The synthetic experiment (simulation.py:250-493) against the FP64 oracle's
composition of the same steps: DISORDER and sequential masks, the reference's
motion generator, corrupted k-space with noise of std max(ksp) 1e-4, the SENSE
baseline and the pyramid.  Small size (32 x 32 x 8, 4 coils), one joint
iteration per level, 2 levels and 20 final CG iterations (the pyramid test's
regime).  Relative L2 1e-4 on k-space and images, 1e-3 (degrees, voxels) on
motion; the written NIfTI and npz files are read back.  The real-data driver
(recon.py:130-172) likewise, from synthetic scanner arrays."""
import os

import numpy as np
import pytest
import torch

from espirit_data import espirit_data
from oracle import motion_oracle as O

pytestmark = pytest.mark.gpu
SHAPE = (32, 32, 8)
NC = 4
TOL = 1e-4


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b))


def _complex_max(x):
    re = x.real.max()
    return np.complex64(complex(float(re), float(x.imag[x.real == re].max())))


@pytest.mark.parametrize("order", ["disorder", "sequential"])
def test_run_experiment_matches_oracle(me, tmp_path, order):
    sim = me.simulation
    gt = sim.phantom(SHAPE, seed=0).numpy()
    mps = sim.coil_maps(SHAPE, NC, seed=1).numpy()
    out = tmp_path / order
    res = sim.run_experiment(gt, mps, tile_size=(4, 4), accel=2, order=order, iters=1, noise=1, out_dir=str(out),
                             sense_iter=20, pyramid_kwargs=dict(n_spatial_levels=2, final_cg_iter=20))
    nshots = 8
    # masks and motion as the reference builds them
    if order == "disorder":
        cfg = O.DisorderConfig(SHAPE[1] - 1, SHAPE[0] - 1, SHAPE[1] // 2, 4, 4, seed=123456)
        ordr, _ = O.make_disorder_order(cfg)
        A = O.disorder_masks(ordr, nshots, SHAPE[:2])
    else:
        y = np.repeat(np.arange(0, SHAPE[0], 2), SHAPE[1])
        x = np.tile(np.arange(SHAPE[1]), SHAPE[0] // 2)
        A = np.zeros((nshots, *SHAPE[:2]), dtype=bool)
        n = y.size // nshots
        for s in range(nshots):
            A[s, y[s * n:(s + 1) * n], x[s * n:(s + 1) * n]] = True
        A = A[..., None]
    assert np.array_equal(res["A"], A) and np.array_equal(res["A_corr"], A)
    t_corr = sim.generate_motion_parameters_new(nshots, low_freq_var=1, high_freq_var=20.0)
    assert np.array_equal(res["t_corr"], t_corr)
    g64, m64 = gt.astype(np.complex128), mps.astype(np.complex128)
    ksp = O.generate_corrupted_kspace(g64, m64, A, t_corr)
    std = _complex_max(ksp.astype(np.complex64)) * np.float32(1e-4)
    ksp = ksp + sim.generate_complex_gaussian_noise(ksp.shape, std=std)
    assert _rel(res["ksp"].cpu().numpy(), ksp) < TOL
    P = 1.0 / (np.sum(np.abs(m64) ** 2, axis=0) + 1e-6)
    unc = O.sense_recon(ksp, m64, P=P, tol=1e-12, max_iter=20)
    assert _rel(res["uncorrected"], unc) < TOL
    corr, t_ref = O.pyramid_reconstruction(ksp, m64, A, nshots, n_spatial_levels=2, n_joint_iters=1,
                                           final_cg_iter=20)
    assert _rel(res["corrected"], corr) < TOL
    d = np.abs(res["t_estimates"] - t_ref)
    assert max(d[:, :3].max(), np.degrees(d[:, 3:]).max()) < 1e-3
    err = np.abs(corr - g64)
    assert abs(res["err_img"] - np.linalg.norm(err)) / np.linalg.norm(err) < TOL
    # files
    data, hdr = me.io.load_nifti(os.path.join(out, "corrected.nii.gz"))
    assert data.dtype == np.float64 and np.array_equal(data, np.abs(res["corrected"]).astype(np.float64))
    z = np.load(os.path.join(out, "results.npz"))
    assert set(z.files) == {"corrected", "uncorrected", "t_estimates", "err_volume", "err_img"}
    assert np.array_equal(z["t_estimates"], res["t_estimates"])
    assert os.path.exists(os.path.join(out, "loss_2_factor.npy"))


def test_run_experiment_argument_checks(me):
    sim = me.simulation
    gt = np.zeros(SHAPE, dtype=np.complex64)
    mps = np.zeros((NC, *SHAPE), dtype=np.complex64)
    with pytest.raises(AssertionError):
        sim.run_experiment(gt, mps, tile_size=(5, 4))
    with pytest.raises(AssertionError):
        sim.run_experiment(gt, mps[:, :16], tile_size=(4, 4))
    with pytest.raises(ValueError):
        sim.run_experiment(gt, mps, tile_size=(4, 4), order="radial")


@pytest.mark.parametrize("sampling", ["disorder", "sequential"])
def test_reconstruct_real_data(me, tmp_path, sampling):
    """Its calibration equals prepare_real_data (bitwise), and the SENSE image
    and the pyramid match the oracle run on the same compressed data and
    counter masks."""
    imsize, nc, nshots = (24, 16, 16), 6, 4
    calib, _, _ = espirit_data(imsize, nc=nc, cal=imsize, seed=9)
    ksp = np.transpose(calib, (3, 0, 1, 2)).astype(np.complex64)
    rng = np.random.default_rng(6)
    noise = (0.01 * (rng.standard_normal((nc, 3000)) + 1j * rng.standard_normal((nc, 3000)))).astype(np.complex64)
    order = rng.permutation(imsize[0] * imsize[1])
    parts, lines = order // imsize[1], order % imsize[1]
    pk = dict(n_spatial_levels=2, final_cg_iter=20)
    res = me.io.reconstruct_real_data(ksp, ksp, noise, lines, parts, nshots, out_dir=str(tmp_path), sampling=sampling,
                                      iters=1, calib_size=12, sense_iter=20, pyramid_kwargs=pk)
    prep = me.io.prepare_real_data(ksp, ksp, noise, lines=lines, partitions=parts, n_shots=nshots, calib_size=12)
    assert torch.equal(res["ksp"], prep["ksp"]) and torch.equal(res["mps"], prep["mps"])
    assert np.array_equal(res["shot_mask"], O.shot_masks_from_counters(lines, parts, nshots, imsize[0], imsize[1]))
    y = res["ksp"].cpu().numpy().astype(np.complex128)
    s = res["mps"].cpu().numpy().astype(np.complex128)
    sense = O.sense_recon(y, s, max_iter=20)
    assert _rel(res["sense"], sense) < TOL
    x_ref, t_ref = O.pyramid_reconstruction(y, s, res["shot_mask"], nshots, n_joint_iters=1, sampling=sampling, **pk)
    assert _rel(res["recon"], x_ref) < TOL
    d = np.abs(res["t_est"] - t_ref)
    assert max(d[:, :3].max(), np.degrees(d[:, 3:]).max()) < 1e-3
    pre, rec = tmp_path / "pre", tmp_path / "recon"
    for name in ("ksp", "ref", "noise", "white_ksp", "mps", "cc_mps", "cc_ksp"):
        assert (pre / f"{name}.npy").exists(), name
    assert np.array_equal(np.load(pre / "cc_ksp.npy"), res["ksp"].cpu().numpy())
    assert np.array_equal(np.load(rec / "joint_recon.npy"), res["recon"])
    assert np.array_equal(np.load(rec / "estimated_transforms.npy"), res["t_est"])
    data, _ = me.io.load_nifti(str(rec / "joint_recon.nii.gz"))
    assert np.array_equal(data, np.abs(res["recon"]).astype(np.float64))
