"""GPU parity of the uncorrected SENSE baseline (sigpy.mri.app.SenseRecon as
called at simulation.py:462-466) against the FP64 oracle restatement:
relative L2 1e-4 on the image."""
import numpy as np
import pytest
import torch

from oracle import motion_oracle as O

pytestmark = pytest.mark.gpu
SHAPE = (16, 12, 8)
NC = 4


def _rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def sdata(me):
    sim = me.simulation
    mps = sim.coil_maps(SHAPE, NC, seed=60).numpy().astype(np.complex128)
    A, _ = sim.disorder_sampling(SHAPE, (4, 4), 2)
    lines = A.any(0)[..., 0]  # union of the shots (centred frame)
    rng = np.random.default_rng(61)
    y = (rng.standard_normal((NC, *SHAPE)) + 1j * rng.standard_normal((NC, *SHAPE))) * lines[None, ..., None]
    P = 1 / (np.sum(np.abs(mps) ** 2, axis=0) + 1e-6)
    return dict(mps=mps, y=y, P=P, lines=lines)


@pytest.mark.parametrize("max_iter,lamda,use_p", [(20, 0.0, True), (8, 0.0, False), (12, 0.05, True)])
def test_sense_recon_matches_oracle(me, sdata, max_iter, lamda, use_p):
    d = sdata
    P = d["P"] if use_p else None
    got = me.sense_recon(d["y"], d["mps"], lamda=lamda, P=P, max_iter=max_iter, tol=1e-12)
    ref = O.sense_recon(d["y"], d["mps"], lamda=lamda, P=P, max_iter=max_iter, tol=1e-12)
    assert _rel(got, ref) < 1e-4


def test_sense_full_weights_with_noise_everywhere(me, sdata):
    """The reference adds noise to the whole k-space array, so rss(y) > 0
    everywhere and SigPy's weights become all ones."""
    d = sdata
    rng = np.random.default_rng(62)
    y = d["y"] + 1e-3 * (rng.standard_normal(d["y"].shape) + 1j * rng.standard_normal(d["y"].shape))
    got = me.sense_recon(y, d["mps"], P=d["P"], max_iter=6)
    ref = O.sense_recon(y, d["mps"], P=d["P"], max_iter=6)
    assert _rel(got, ref) < 1e-4


def test_sense_zero_data_and_bad_weights(me, sdata):
    d = sdata
    w = np.broadcast_to(d["lines"][..., None], SHAPE).astype(np.float64)
    app = me.SenseRecon(np.zeros_like(d["y"]), d["mps"], weights=w, max_iter=5)
    x = app.run()
    assert app.iterations == 0 and float(x.abs().max()) == 0.0
    with pytest.raises(NotImplementedError):
        me.SenseRecon(d["y"], d["mps"], weights=0.5 * w)
