"""GPU parity of the calibration preprocessing (noise_whiten.py:5-14,
coil_compression.py:6-46) against the FP64 oracle, and of the Hermitian
eigensolver against numpy.linalg.eigh.

Eigenvector phases are not unique: LAPACK's SVD and the Jacobi rotations
choose different unit phases per virtual coil, which leave the compressed
reconstruction unchanged.  Coil compression is therefore compared through
phase-invariant quantities (kept count, singular values, the projector
A^H A) and, per virtual coil, after aligning the one unit phase."""
import numpy as np
import pytest
import torch

from oracle import motion_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n", [1, 7, 32, 64])
def test_hermitian_eig_matches_eigh(me, n):
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, 3 * n)) + 1j * rng.standard_normal((n, 3 * n))
    H = X @ X.conj().T
    dev = torch.device("cuda:0")
    Hd = torch.view_as_real(torch.as_tensor(H, device=dev)).contiguous()
    lam = torch.empty(n, dtype=torch.float64, device=dev)
    V = torch.empty((n, n, 2), dtype=torch.float64, device=dev)
    me._lib.call("me_hermitian_eig", me._lib.ptr(Hd), n, 60, me._lib.ptr(lam), me._lib.ptr(V), me._lib.stream(dev))
    lam = lam.cpu().numpy()
    V = torch.view_as_complex(V).cpu().numpy()
    ref = np.linalg.eigvalsh(H)[::-1]
    assert np.max(np.abs(lam - ref)) <= 1e-12 * ref[0]
    assert np.linalg.norm(V @ np.diag(lam) @ V.conj().T - H) <= 1e-12 * np.linalg.norm(H)
    assert np.linalg.norm(V.conj().T @ V - np.eye(n)) <= 1e-12 * n


def test_noise_whiten_matches_oracle(me):
    rng = np.random.default_rng(72)
    C = 12
    mix = rng.standard_normal((C, C)) + 1j * rng.standard_normal((C, C))
    noise = (mix @ (rng.standard_normal((C, 5000)) + 1j * rng.standard_normal((C, 5000)))).astype(np.complex64)
    ksp = (rng.standard_normal((C, 16, 12, 8)) + 1j * rng.standard_normal((C, 16, 12, 8))).astype(np.complex64)
    got, W = me.noise_whiten(ksp, noise)
    ref, Wref = O.noise_whiten(ksp, noise)
    assert _rel(W, Wref) < 1e-9
    assert _rel(got, ref) < 1e-5


@pytest.mark.parametrize("perc", [0.95, 0.99])
def test_coil_compression_matches_oracle(me, perc):
    sim = me.simulation
    shape = (24, 20, 12)
    mps = sim.coil_maps(shape, 16, seed=73).numpy()
    rng = np.random.default_rng(74)
    ksp = (rng.standard_normal((16, *shape)) + 1j * rng.standard_normal((16, *shape))).astype(np.complex64)
    cc, cck = me.coil_compression(mps, ksp, perc=perc)
    rc, rck, A, S = O.coil_compression(mps, ksp, perc=perc)
    assert cc.shape == rc.shape and cck.shape == rck.shape and cc.shape[0] >= 1
    lam = me.coil_compression.singular_values
    assert np.max(np.abs(lam - S)) <= 1e-6 * S[0]
    Ag = me.coil_compression.matrix.cpu().numpy().astype(np.complex128)
    assert _rel(Ag.conj().T @ Ag, A.conj().T @ A) < 1e-6
    ccn, cckn = cc.cpu().numpy(), cck.cpu().numpy()
    for i in range(rc.shape[0]):  # align the unit phase of virtual coil i
        ph = np.vdot(ccn[i], rc[i])
        ph /= abs(ph)
        assert _rel(ph * ccn[i], rc[i]) < 1e-5
        assert _rel(ph * cckn[i], rck[i]) < 1e-5
