"""GPU metrics and the real-data calibration chain (simulation.py:87-97,
476-477; recon.py:130-165) against the oracle.

The gradient-entropy histogram counts are integers and must equal NumPy's
bit for bit (the kernel follows NumPy's float32 uniform-bin rule, edges
included); the entropy then agrees to FP64 rounding."""
import numpy as np
import pytest
import torch

from espirit_data import espirit_data
from oracle import motion_oracle as O

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


@pytest.mark.parametrize("shape", [(16, 12, 8), (33, 20, 17), (64, 64, 48)])
def test_gradient_entropy_matches_numpy(me, shape):
    ph = me.simulation.phantom(shape, seed=3, device=DEV)
    vol = ph.abs().float().cpu().numpy()
    ent, counts = O.gradient_entropy(vol)
    got = me.metrics.gradient_entropy(vol)
    g = [np.diff(vol, axis=a, append=np.take(vol, [-1], axis=a)) for a in range(3)]
    assert me.metrics.gradient_entropy.max == float(np.sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]).max())
    assert np.array_equal(me.metrics.gradient_entropy.counts.cpu().numpy(), counts)
    assert abs(got - ent) <= 1e-12 * abs(ent)


def test_gradient_entropy_special_volumes(me):
    for vol in (np.full((6, 5, 4), 3.0, np.float32),
                np.broadcast_to((0.5 * np.arange(6, dtype=np.float32))[None, :, None], (4, 6, 5)).copy()):
        ent, counts = O.gradient_entropy(vol)
        got = me.metrics.gradient_entropy(torch.as_tensor(vol, device=DEV))
        assert np.array_equal(me.metrics.gradient_entropy.counts.cpu().numpy(), counts)
        assert abs(got - ent) <= 1e-12 * abs(ent)


def test_error_volume_matches_oracle(me):
    rng = np.random.default_rng(4)
    a = (rng.standard_normal((20, 18, 9)) + 1j * rng.standard_normal((20, 18, 9))).astype(np.complex64)
    b = (rng.standard_normal((20, 18, 9)) + 1j * rng.standard_normal((20, 18, 9))).astype(np.complex64)
    e, n = me.metrics.error_volume(a, b)
    er, nr = O.error_volume(a.astype(np.complex128), b.astype(np.complex128))
    assert np.max(np.abs(e.cpu().numpy() - er)) <= 1e-6 * er.max()
    assert abs(n - nr) <= 1e-6 * nr


def test_save_results_npz(me, tmp_path):
    rng = np.random.default_rng(5)
    c = (rng.standard_normal((8, 6, 4)) + 1j * rng.standard_normal((8, 6, 4))).astype(np.complex64)
    g = (rng.standard_normal((8, 6, 4)) + 1j * rng.standard_normal((8, 6, 4))).astype(np.complex64)
    err = me.io.save_results(str(tmp_path / "results"), c, 0.5 * c, np.zeros((3, 6)), g)
    z = np.load(tmp_path / "results.npz")
    assert set(z.files) == {"corrected", "uncorrected", "t_estimates", "err_volume", "err_img"}
    er, nr = O.error_volume(c, g)
    assert abs(float(z["err_img"]) - nr) <= 1e-6 * nr and abs(err - nr) <= 1e-6 * nr
    assert np.allclose(z["err_volume"], er, rtol=1e-5, atol=1e-6)


def test_prepare_real_data_chain(me):
    """The chain equals its parts composed by hand, and the shot masks equal
    the oracle's; the parts themselves are checked against the oracle in
    test_gpu_calibration and test_gpu_espirit."""
    imsize, nc = (24, 20, 16), 6
    calib, maps, obj = espirit_data(imsize, nc=nc, cal=imsize, seed=9)
    ksp = np.transpose(calib, (3, 0, 1, 2)).astype(np.complex64)  # (coils, par, lin, col)
    rng = np.random.default_rng(6)
    noise = (0.01 * (rng.standard_normal((nc, 3000)) + 1j * rng.standard_normal((nc, 3000)))).astype(np.complex64)
    order = rng.permutation(imsize[0] * imsize[1])
    parts, lines = order // imsize[1], order % imsize[1]
    out = me.io.prepare_real_data(ksp, ksp, noise, lines=lines, partitions=parts, n_shots=5, calib_size=12,
                                  kernel_size=(3, 3, 3))
    white, W = me.noise_whiten(ksp, noise)
    ref = me.io._centred_resize(torch.as_tensor(ksp, device=DEV), (nc, 12, 12, 12))
    mps = me.rx_espirit_3d(ref.permute(1, 2, 3, 0).contiguous(), imsize, kernel_size=(3, 3, 3))
    cc_mps, cc_ksp = me.coil_compression(mps.permute(3, 0, 1, 2).contiguous(), white, perc=0.95)
    assert torch.equal(out["mps"], cc_mps) and torch.equal(out["ksp"], cc_ksp) and torch.equal(out["W"], W)
    assert np.array_equal(out["shot_mask"], O.shot_masks_from_counters(lines, parts, 5, imsize[0], imsize[1]))
    assert out["mps"].shape[0] >= 1 and torch.isfinite(torch.view_as_real(out["mps"])).all()
