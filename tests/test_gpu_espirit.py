"""GPU parity of ESPIRiT (espirit.py:9-239) against the FP64 oracle, and of
its building blocks (calibration matrix, FP64 Gram, parallel Jacobi
eigensolver) against NumPy.

The maps are compared voxel by voxel after the reference's own phase
normalisation (coil 0 real and non-negative), which makes the principal
eigenvector unique; voxels whose mask value sqrt(lambda_max) * prod(N) /
sqrt(prod(kernel)) lies within 1e-4 (relative) of the threshold are left
out of the mask comparison, where FP32 and FP64 may round either way."""
import ctypes as C

import numpy as np
import pytest
import torch

from espirit_data import espirit_data
from oracle import motion_oracle as O

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _ints(v):
    return (C.c_int * len(v))(*[int(x) for x in v])


def test_calibration_matrix_matches_oracle(me):
    rng = np.random.default_rng(3)
    calib = (rng.standard_normal((7, 6, 5, 3)) + 1j * rng.standard_normal((7, 6, 5, 3))).astype(np.complex64)
    kernel = (3, 2, 4)
    ref = O.fold_rx(O.hankel_fwd(calib.astype(np.complex128), kernel))
    dims, kern = _ints(calib.shape), _ints(kernel)
    n = me._lib.size("me_espirit_hankel_size", dims, kern)
    assert n == ref.size
    H = torch.empty(n, dtype=torch.complex64, device=DEV)
    me._lib.call("me_espirit_calibration_matrix", me._lib.ptr(torch.as_tensor(calib, device=DEV)), dims, kern,
                 me._lib.ptr(H), me._lib.stream(DEV))
    assert np.array_equal(H.cpu().numpy().reshape(ref.shape), ref.astype(np.complex64))
    assert me._lib.size("me_espirit_hankel_size", dims, _ints((8, 2, 2))) == 0


@pytest.mark.parametrize("m,n", [(1, 5), (33, 70), (100, 40)])
def test_gram_rows_matches_numpy(me, m, n):
    rng = np.random.default_rng(m)
    X = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(np.complex64)
    G = torch.empty((m, m, 2), dtype=torch.float64, device=DEV)
    me._lib.call("me_gram_rows", me._lib.ptr(torch.as_tensor(X, device=DEV)), m, n, me._lib.ptr(G),
                 me._lib.stream(DEV))
    Xd = X.astype(np.complex128)
    ref = Xd @ Xd.conj().T
    got = torch.view_as_complex(G).cpu().numpy()
    assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)


def _jacobi(me, A, sweeps=40, tol=1e-14):
    m = A.shape[0]
    G = torch.view_as_real(torch.as_tensor(A, device=DEV)).contiguous()
    V = torch.empty((m, m, 2), dtype=torch.float64, device=DEV)
    lam = torch.empty(m, dtype=torch.float64, device=DEV)
    work = torch.empty(me._lib.size("me_jacobi_workspace_size", m), dtype=torch.float64, device=DEV)
    done = C.c_int(0)
    me._lib.call("me_jacobi_eig", me._lib.ptr(G), me._lib.ptr(V), m, sweeps, tol, me._lib.ptr(lam),
                 me._lib.ptr(work), C.byref(done), me._lib.stream(DEV))
    return lam.cpu().numpy(), torch.view_as_complex(V).cpu().numpy(), done.value


@pytest.mark.parametrize("m", [5, 64, 127])
def test_jacobi_global_v_updates(me, m):
    """The path for matrices above 8192 rows (V rotated in global memory in
    every round instead of 32 rounds at a time in LDS), forced at small m:
    the same rotations in the same order, so the same eigenpairs."""
    rng = np.random.default_rng(30 + m)
    X = rng.standard_normal((m, m + 2)) + 1j * rng.standard_normal((m, m + 2))
    A = X @ X.conj().T
    lam0, V0, sw0 = _jacobi(me, A)
    prev = me._lib.size("me_jacobi_set_lds_rows", 0)
    try:
        lam1, V1, sw1 = _jacobi(me, A)
    finally:
        me._lib.size("me_jacobi_set_lds_rows", prev)
    assert prev == 8192 and sw1 == sw0
    assert np.array_equal(lam1, lam0)
    assert np.linalg.norm(V1 - V0) <= 1e-13 * np.sqrt(m)
    assert np.linalg.norm(V1 @ np.diag(lam1) @ V1.conj().T - A) <= 1e-12 * np.linalg.norm(A)


@pytest.mark.parametrize("m", [1, 2, 5, 64, 127])
def test_jacobi_eig_matches_eigh(me, m):
    rng = np.random.default_rng(10 + m)
    X = rng.standard_normal((m, m + 3)) + 1j * rng.standard_normal((m, m + 3))
    A = X @ X.conj().T
    lam, V, sweeps = _jacobi(me, A)
    ref = np.linalg.eigvalsh(A)
    assert np.max(np.abs(np.sort(lam) - ref)) <= 1e-12 * ref[-1]
    assert np.linalg.norm(V @ np.diag(lam) @ V.conj().T - A) <= 1e-12 * np.linalg.norm(A)
    assert np.linalg.norm(V.conj().T @ V - np.eye(m)) <= 1e-12 * m
    assert sweeps < 40


def test_jacobi_eig_rank_deficient(me):
    """Rank 6 of 40 (the shape of a calibration Gram matrix with more rows
    than blocks): the zero eigenvalues come out at the round-off level and
    the projector onto the leading 6 equals X's column-space projector."""
    rng = np.random.default_rng(4)
    X = rng.standard_normal((40, 6)) + 1j * rng.standard_normal((40, 6))
    A = X @ X.conj().T
    lam, V, _ = _jacobi(me, A)
    order = np.argsort(-lam)
    assert np.all(np.abs(lam[order[6:]]) <= 1e-12 * lam[order[0]])
    U = V[:, order[:6]]
    Q, _ = np.linalg.qr(X)
    assert np.linalg.norm(U @ U.conj().T - Q @ Q.conj().T) <= 1e-10


def _compare_maps(sens, ref, info, mask_thresh):
    got = sens.cpu().numpy().astype(np.complex128)
    val = info["mask_value"]
    border = np.abs(val - mask_thresh) <= 1e-4 * max(mask_thresh, 1e-30)
    m_ref = np.any(ref != 0, axis=-1)
    m_got = np.any(got != 0, axis=-1)
    assert np.array_equal(m_ref[~border], m_got[~border])
    both = m_ref & m_got
    assert both.sum() > 0
    err = np.linalg.norm(got[both] - ref[both]) / np.linalg.norm(ref[both])
    assert err <= 1e-4, err
    # unit vectors, coil 0 real and non-negative
    assert np.allclose(np.linalg.norm(got[both], axis=-1), 1.0, atol=1e-5)
    assert np.all(np.abs(got[both][:, 0].imag) <= 1e-6) and np.all(got[both][:, 0].real >= 0)


@pytest.mark.parametrize("imsize,nc,cal,kernel,eig_thresh,mask_thresh", [
    ((24, 20, 16), 4, (12, 12, 10), (3, 3, 3), 0.02, 0.99),
    ((25, 19, 15), 4, (11, 12, 9), (3, 2, 3), 0.02, 0.9),
    ((32, 28, 20), 8, (14, 14, 12), (5, 5, 5), 0.02, 0.99),
    ((24, 20, 16), 6, (12, 12, 10), (3, 3, 3), 20, 0.5),
])
def test_espirit_matches_oracle(me, imsize, nc, cal, kernel, eig_thresh, mask_thresh):
    calib, _, _ = espirit_data(imsize, nc=nc, cal=cal, seed=sum(imsize) + nc)
    calib = calib.astype(np.complex64)
    info = {}
    ref = O.rx_espirit_3d(calib, imsize, kernel_size=kernel, eig_thresh=eig_thresh, mask_thresh=mask_thresh,
                          info=info)
    sens = me.rx_espirit_3d(calib, imsize, kernel_size=kernel, eig_thresh=eig_thresh, mask_thresh=mask_thresh)
    assert me.rx_espirit_3d.kept == info["kept"]
    S = info["singular_values"]
    sv = me.rx_espirit_3d.singular_values
    k = int(np.sum(S > 1e-3 * S[0]))  # below, the Jacobi tolerance dominates
    assert np.max(np.abs(sv[:k] - S[:k])) <= 1e-6 * S[0]
    _compare_maps(sens, ref, info, mask_thresh)


@pytest.mark.parametrize("opa,opb,m,n,k", [(0, 0, 37, 45, 29), (1, 0, 33, 17, 70), (0, 1, 5, 64, 40),
                                            (1, 1, 64, 31, 33)])
def test_zgemm_matches_numpy(me, opa, opb, m, n, k):
    rng = np.random.default_rng(m + n + k)
    a_shape = (m, k) if opa == 0 else (k, m)
    b_shape = (k, n) if opb == 0 else (n, k)
    A = rng.standard_normal(a_shape) + 1j * rng.standard_normal(a_shape)
    B = rng.standard_normal(b_shape) + 1j * rng.standard_normal(b_shape)
    Ad = torch.view_as_real(torch.as_tensor(A, device=DEV)).contiguous()
    Bd = torch.view_as_real(torch.as_tensor(B, device=DEV)).contiguous()
    Cd = torch.empty((m, n, 2), dtype=torch.float64, device=DEV)
    me._lib.call("me_zgemm", opa, opb, m, n, k, me._lib.ptr(Ad), me._lib.ptr(Bd), me._lib.ptr(Cd),
                 me._lib.stream(DEV))
    ref = (A if opa == 0 else A.conj().T) @ (B if opb == 0 else B.conj().T)
    got = torch.view_as_complex(Cd).cpu().numpy()
    assert np.linalg.norm(got - ref) <= 1e-14 * np.linalg.norm(ref) * np.sqrt(k)


def test_random_normal_moments(me):
    n = 1 << 18
    x = torch.empty((n, 2), dtype=torch.float64, device=DEV)
    me._lib.call("me_random_normal", me._lib.ptr(x), n, 12345, me._lib.stream(DEV))
    v = x.cpu().numpy()
    assert abs(v.mean()) < 0.01 and abs(v.std() - 1.0) < 0.01
    y = torch.empty_like(x)
    me._lib.call("me_random_normal", me._lib.ptr(y), n, 12345, me._lib.stream(DEV))
    assert torch.equal(x, y)


@pytest.mark.parametrize("method", ["jacobi", "subspace"])
def test_espirit_many_coils_both_routes(me, method):
    """64 coils with a 3x3x3 kernel (m = 1728 rows): the kept subspace by full
    Jacobi and by subspace iteration with Rayleigh-Ritz, and the per-row
    coil-matrix table (64 * 64 * 5 complex64) beyond LDS, so the tables are
    built in global memory in chunks of rows; same parity bar either way."""
    imsize, nc = (20, 16, 12), 64
    calib, _, _ = espirit_data(imsize, nc=nc, cal=(10, 10, 8), seed=77)
    calib = calib.astype(np.complex64)
    info = {}
    ref = O.rx_espirit_3d(calib, imsize, kernel_size=(3, 3, 3), info=info)
    sens = me.rx_espirit_3d(calib, imsize, kernel_size=(3, 3, 3), method=method)
    assert me.rx_espirit_3d.method == method
    assert me.rx_espirit_3d.kept == info["kept"]
    S, sv = info["singular_values"], me.rx_espirit_3d.singular_values
    # the Ritz values are accurate where the block has converged: down to
    # and past the keep threshold 0.02 * s0
    kk = int(np.sum(S > 0.01 * S[0]))
    assert np.max(np.abs(sv[:kk] - S[:kk])) <= 1e-6 * S[0]
    _compare_maps(sens, ref, info, 0.99)


def test_espirit_nothing_kept_gives_zero_maps(me):
    calib = np.zeros((8, 8, 8, 3), dtype=np.complex64)
    sens = me.rx_espirit_3d(calib, (10, 10, 10), kernel_size=(3, 3, 3))
    assert sens.shape == (10, 10, 10, 3) and not torch.any(sens != 0)
    assert me.rx_espirit_3d.kept == 0


def test_espirit_rejects_oversized_kernel(me):
    calib = np.ones((4, 4, 4, 2), dtype=np.complex64)
    with pytest.raises(ValueError):
        me.rx_espirit_3d(calib, (8, 8, 8), kernel_size=(5, 5, 5))
