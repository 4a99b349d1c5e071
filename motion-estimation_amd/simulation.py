"""Synthetic inputs: DISORDER sampling (simulation.py:192-245, 383-410), the
motion-corrupted forward model (simulation.py:65-85), and the seeded
phantom / coil / motion / noise generators the benchmark configurations
are defined with.  Also the reference's own NumPy generators
(``generate_motion_parameters_new``, ``generate_complex_gaussian_noise``,
``sample_disorder``; same draws, same outputs) and its synthetic experiment
(simulation.py:250-493) as :func:`run_experiment`.

DISORDER is computed by the library's host C code (bit-exact with the
reference's stdlib ``random``); the corrupted k-space is synthesised on the GPU
with the same encoding kernels as the reconstruction.  The generators
evaluate volumes with torch on the target device from numpy-seeded
parameters, so every configuration is reproducible from its seeds.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .encoding import ShotPlan, permute, plane_strides, shift_offsets, std_strides, to_plane
from .transform import Geometry
from .utils import as_c64, as_f64, compute_transform_grids_voxel


@dataclass
class DisorderConfig:
    """simulation.py:192-199."""
    max_line_number: int
    max_partition_number: int
    pat_lines_to_measure: int
    block_lin: int
    block_par: int
    seed: int = 123456


def make_disorder_order(cfg: DisorderConfig):
    """simulation.py:201-245: list of (line, partition) and the padded (L, P)."""
    cnt = C.c_longlong(0)
    L = C.c_int(0)
    P = C.c_int(0)
    lib = _lib.load()
    lib.me_disorder_order(cfg.max_line_number, cfg.max_partition_number, cfg.pat_lines_to_measure, cfg.block_lin,
                          cfg.block_par, cfg.seed, None, None, 0, C.byref(cnt), C.byref(L), C.byref(P))
    n = cnt.value
    lines = np.zeros(max(1, n), dtype=np.int32)
    parts = np.zeros(max(1, n), dtype=np.int32)
    _lib.call("me_disorder_order", cfg.max_line_number, cfg.max_partition_number, cfg.pat_lines_to_measure,
              cfg.block_lin, cfg.block_par, cfg.seed, lines.ctypes.data_as(C.c_void_p),
              parts.ctypes.data_as(C.c_void_p), n, C.byref(cnt), C.byref(L), C.byref(P))
    order = list(zip(lines[:n].tolist(), parts[:n].tolist()))
    return order, (L.value, P.value)


def disorder_masks(order, nshots, K):
    """simulation.py:396-402 and 437: shot s takes the contiguous chunk
    order[s*n:(s+1)*n], n = len(order)//nshots, mask[s, partition, line];
    leftover samples are dropped.  Returns (nshots, K0, K1, 1) bool."""
    order = np.asarray(order, dtype=np.int64).reshape(-1, 2)
    A = np.zeros((nshots, *K), dtype=bool)
    sps = order.shape[0] // nshots
    for s in range(nshots):
        A[s, order[s * sps:(s + 1) * sps, 1], order[s * sps:(s + 1) * sps, 0]] = True
    return A[..., None]


def disorder_sampling(shape, tile, accel, nshots=None, seed=123456):
    """DISORDER masks for an (N0, N1, N2) image as simulation.py:383-402 does:
    tile = (partitions, lines); default nshots = prod(tile)//accel
    (simulation.py:372)."""
    cfg = DisorderConfig(max_line_number=shape[1] - 1, max_partition_number=shape[0] - 1,
                         pat_lines_to_measure=shape[1] // accel, block_lin=tile[1], block_par=tile[0], seed=seed)
    order, _ = make_disorder_order(cfg)
    ns = nshots if nshots is not None else (tile[0] * tile[1]) // accel
    return disorder_masks(order, ns, tuple(shape[:2])), order


def generate_corrupted_kspace(gt_img, mps, A, transforms, batch=None):
    """simulation.py:65-85 on the GPU: centred image/maps/masks in, centred
    k-space out, ksp = fftshift(sum_s A_s fftn(S T_s ifftshift(gt)))."""
    dev = _lib.require_gpu()
    gt = as_c64(gt_img, dev)
    maps = as_c64(mps, dev)
    shape = tuple(gt.shape)
    n0, n1, n2 = shape
    vol = n0 * n1 * n2
    ncoils = maps.shape[0]
    kgrid, rkgrid = compute_transform_grids_voxel(shape, [1, 1, 1])
    geom = Geometry.from_grids(kgrid, rkgrid, dev)
    xp = to_plane(gt, shape, "i")
    mp = to_plane(maps, shape, "i", ncoils=ncoils)
    mask = np.fft.ifftshift(np.asarray(A.cpu().numpy() if isinstance(A, torch.Tensor) else A), axes=(1, 2))
    plan = ShotPlan(mask, dev)
    theta = as_f64(transforms, dev).reshape(-1, 6)
    S = plan.nshots
    B = batch or max(1, min(S, int(4 * 2**30) // (vol * 8)))
    U = torch.empty((B, vol), dtype=torch.complex64, device=dev)
    Y = torch.empty(max(1, plan.total * n2 * ncoils), dtype=torch.complex64, device=dev)
    s = _lib.stream(dev)
    for s0 in range(0, S, B):
        s1 = min(S, s0 + B)
        coef = torch.empty((s1 - s0, 16), dtype=torch.float64, device=dev)
        _lib.call("me_rigid_transform", _lib.ptr(xp), 0, 1, _lib.ptr(U), vol, None, 0, s1 - s0, geom.n_arg(),
                  geom.st_arg(True), *geom.r_args(), _lib.ptr(theta[s0:s1]), 6, _lib.ptr(coef), 0, 0, None, s)
        maxc = int(plan.cnt_h[s0:s1].max())
        _lib.call("me_encode_forward", _lib.ptr(U), vol, _lib.ptr(mp), ncoils, s1 - s0, _lib.int3(shape),
                  _lib.ptr(plan.sk0), _lib.ptr(plan.sk1), _lib.ptr(plan.off[s0:s1]), _lib.ptr(plan.cnt[s0:s1]), maxc,
                  _lib.ptr(Y), None, None, None, 0, s)
    del U
    kh = torch.zeros((ncoils, *shape), dtype=torch.complex64, device=dev)
    _lib.call("me_scatter_hybrid", _lib.ptr(Y), ncoils, _lib.int3(shape), _lib.ptr(plan.sk0), _lib.ptr(plan.sk1),
              _lib.ptr(plan.off), _lib.ptr(plan.cnt), S, plan.total, _lib.ptr(kh), s)
    _lib.call("me_fft_axis", _lib.ptr(kh), vol, _lib.ptr(kh), vol, ncoils, _lib.int3(shape),
              _lib.ll3((n1 * n2, n2, 1)), 2, 0, s)
    out = torch.empty_like(kh)
    dims = (ncoils, *shape)
    permute(kh, std_strides(0, shape), dims, out, std_strides(0, shape), dims, shift_offsets(shape, "f"))
    return out


# ---------------------------------------------------------------------------
# seeded synthetic generators (configuration inputs; not reference functions)
# ---------------------------------------------------------------------------
def _centered_coords(shape, device):
    return [torch.arange(n, device=device, dtype=torch.float32) - n // 2 for n in shape]


def _rotmat(a, b, c):
    ca, sa, cb, sb, cc, sc = math.cos(a), math.sin(a), math.cos(b), math.sin(b), math.cos(c), math.sin(c)
    Rx = np.array([[1, 0, 0], [0, ca, -sa], [0, sa, ca]])
    Ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    Rz = np.array([[cc, -sc, 0], [sc, cc, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def phantom(shape, n_ellipsoids=10, seed=0, device=None):
    """Sum of ``n_ellipsoids`` random ellipsoids (centres within 60% of the
    half field of view, semi-axes U[0.1, 0.4]*N, random orientation,
    magnitudes U[0.1, 1]) times exp(i*phi), phi a random polynomial of degree
    <= 2 scaled to +-pi.  Centred frame, complex64."""
    device = device or torch.device("cpu")
    rng = np.random.default_rng(seed)
    g = torch.meshgrid(*_centered_coords(shape, device), indexing="ij")
    half = np.array(shape, dtype=np.float64) / 2
    img = torch.zeros(shape, dtype=torch.float32, device=device)
    for _ in range(n_ellipsoids):
        c = rng.uniform(-0.6, 0.6, 3) * half
        ax = rng.uniform(0.1, 0.4, 3) * np.array(shape)
        R = _rotmat(*rng.uniform(0, 2 * np.pi, 3))
        mag = rng.uniform(0.1, 1.0)
        d = [g[i] - float(c[i]) for i in range(3)]
        q = torch.zeros(shape, dtype=torch.float32, device=device)
        for row in range(3):
            u = float(R[0, row]) * d[0] + float(R[1, row]) * d[1] + float(R[2, row]) * d[2]
            q += (u / float(ax[row])) ** 2
        img += mag * (q <= 1.0).float()
    # smooth phase: degree-2 polynomial in normalised coordinates
    coef = rng.uniform(-1, 1, 10)
    xn = [g[i] / float(max(1.0, half[i])) for i in range(3)]
    terms = [torch.ones_like(xn[0]), xn[0], xn[1], xn[2], xn[0] ** 2, xn[1] ** 2, xn[2] ** 2, xn[0] * xn[1],
             xn[0] * xn[2], xn[1] * xn[2]]
    phi = sum(float(a) * t for a, t in zip(coef, terms))
    phi = phi - phi.mean()
    phi = phi * (math.pi / float(phi.abs().max().clamp_min(1e-12)))
    return (img * torch.exp(1j * phi)).to(torch.complex64)


def coil_maps(shape, ncoils, seed=1, device=None):
    """``ncoils`` smooth Gaussian-falloff maps on a ring around the readout
    axis 2 (alternating between two z positions when ncoils > 8), each with a
    small linear phase, normalised so sum_c |S_c|^2 = 1.  Centred frame."""
    device = device or torch.device("cpu")
    rng = np.random.default_rng(seed)
    g = torch.meshgrid(*_centered_coords(shape, device), indexing="ij")
    n = np.array(shape, dtype=np.float64)
    maps = []
    for c in range(ncoils):
        ang = 2 * np.pi * c / ncoils
        zc = 0.0 if ncoils <= 8 else (0.25 if c % 2 else -0.25) * n[2]
        ctr = np.array([0.6 * n[0] / 2 * np.cos(ang), 0.6 * n[1] / 2 * np.sin(ang), zc])
        sig = 0.45 * n
        d2 = sum(((g[i] - float(ctr[i])) / float(sig[i])) ** 2 for i in range(3))
        slope = rng.uniform(-1, 1, 3) * np.pi / n
        ph = sum(float(slope[i]) * g[i] for i in range(3))
        maps.append(torch.exp(-0.5 * d2) * torch.exp(1j * ph))
    S = torch.stack(maps)
    S = S / torch.sqrt((S.abs() ** 2).sum(0, keepdim=True))
    return S.to(torch.complex64)


def motion_uniform(nshots, rot_deg=2.0, trans_vox=2.0, seed=2):
    """Rotations U[-rot_deg, rot_deg] (radians out), translations U[-t, t]."""
    rng = np.random.default_rng(seed)
    t = np.zeros((nshots, 6))
    t[:, :3] = rng.uniform(-trans_vox, trans_vox, (nshots, 3))
    t[:, 3:] = np.radians(rng.uniform(-rot_deg, rot_deg, (nshots, 3)))
    return t


def motion_random_walk(nshots, rot_step_deg=0.5, trans_step_vox=0.5, seed=2):
    """Per-shot random walk of the six parameters."""
    rng = np.random.default_rng(seed)
    steps = np.concatenate([rng.normal(0, trans_step_vox, (nshots, 3)),
                            np.radians(rng.normal(0, rot_step_deg, (nshots, 3)))], axis=1)
    return np.cumsum(steps, axis=0)


def motion_smooth_cubic(nstates, amp_deg=5.0, amp_vox=5.0, seed=2):
    """Smooth cubic trajectories over the acquisition, amplitude up to
    +-amp per parameter, sampled at ``nstates`` motion states."""
    rng = np.random.default_rng(seed)
    tt = np.linspace(0.0, 1.0, nstates)
    out = np.zeros((nstates, 6))
    for p in range(6):
        c = rng.uniform(-1, 1, 4)
        v = c[0] + c[1] * tt + c[2] * tt ** 2 + c[3] * tt ** 3
        v = v - v[0]
        v = v / max(1e-12, np.abs(v).max())
        amp = amp_vox if p < 3 else np.radians(amp_deg)
        out[:, p] = v * amp * rng.uniform(0.5, 1.0)
    return out


def add_noise(ksp, snr_db, sampled, seed=3):
    """Complex Gaussian noise over the whole k-space with variance set so
    that mean |y|^2 over the sampled entries / E|n|^2 = 10^(snr/10)."""
    sm = torch.as_tensor(np.asarray(sampled, dtype=bool).reshape(1, ksp.shape[1], ksp.shape[2], 1), device=ksp.device)
    p_sig = float(((ksp.abs() ** 2) * sm).sum() / (sm.sum() * ksp.shape[0] * ksp.shape[3]))
    sigma = math.sqrt(p_sig / 10 ** (snr_db / 10) / 2)
    gen = torch.Generator(device=ksp.device)
    gen.manual_seed(seed)
    noise = torch.randn(ksp.shape, generator=gen, device=ksp.device, dtype=torch.float32) * sigma
    ksp = ksp + noise
    noise = torch.randn(ksp.shape, generator=gen, device=ksp.device, dtype=torch.float32) * sigma
    return (ksp + 1j * noise).to(torch.complex64)


# ------------------------------------------------ the reference's generators
def generate_complex_gaussian_noise(shape, std=1.0, seed=42):
    """simulation.py:99-102 with xp = numpy: std * N1 + 1j * std * N2 from
    one NumPy generator (float32 draws, N1 first), complex64.  ``std`` may be
    complex, as the experiment passes max(ksp) * 1e-4 * no.  The reference's
    CuPy call (xp = cupy) draws from CuPy's generator instead, whose stream
    is not reproduced here."""
    rng = np.random.default_rng(seed)
    n = std * rng.standard_normal(shape, dtype=np.float32) + 1j * std * rng.standard_normal(shape, dtype=np.float32)
    return n.astype(np.complex64)


def generate_motion_parameters_new(num_states, low_freq_var=0.1, high_freq_var=5.0, spike_prob=0.02, seed=123456789,
                                   rot_limit_deg=20.0):
    """simulation.py:104-174: per parameter a linear drift from 0 to a
    Gaussian end point, a smooth AR(1) wobble (rho = 0.98, stationary
    variance 0.3 / 0.2 x low_freq_var for translations / rotations) and one
    or two Gaussian bumps (count round(spike_prob * T) clamped to [1, 2],
    centres U[0.15, 0.85], widths U[0.06, 0.14] x T, amplitudes scaled from
    high_freq_var); rotations (degrees) clipped to +-rot_limit_deg, then
    radians.  Columns (trans x3 in voxels, rot x3 in radians).  The draws come
    from one NumPy generator in the reference's order (drift ends, wobble
    innovations, bump centres, widths, amplitudes), so the output is identical
    for a given seed."""
    rng = np.random.default_rng(seed)
    T = int(num_states)
    tt = np.linspace(0.0, 1.0, T)
    end_trans = rng.normal(0, np.sqrt(low_freq_var), size=3)
    end_rot = rng.normal(0, 0.6 * np.sqrt(low_freq_var), size=3)
    rho = 0.98

    def wobble(target_var):
        sd = np.sqrt(max(1e-12, target_var) * (1 - rho ** 2))
        w = np.zeros((T, 3))
        for i in range(1, T):  # one draw of 3 per step, in step order
            w[i] = rho * w[i - 1] + rng.normal(0, sd, size=3)
        return w

    wob_trans = wobble(0.3 * low_freq_var)
    wob_rot = wobble(0.2 * low_freq_var)
    nb = int(np.clip(max(1, int(round(spike_prob * T))), 1, 2))
    centre = rng.uniform(0.15, 0.85, size=nb)
    width = rng.uniform(0.06, 0.14, size=nb) * T
    amp_trans = rng.normal(0, 0.15 * np.sqrt(high_freq_var), size=(nb, 3))
    amp_rot = rng.normal(0, 0.08 * np.sqrt(high_freq_var), size=(nb, 3))
    bump_trans = np.zeros((T, 3))
    bump_rot = np.zeros((T, 3))
    pos = np.arange(T)
    for k in range(nb):
        g = np.exp(-0.5 * ((pos - centre[k] * T) / (width[k] + 1e-9)) ** 2)[:, None]
        bump_trans += g * amp_trans[k][None, :]
        bump_rot += g * amp_rot[k][None, :]
    trans = tt[:, None] * end_trans[None, :] + wob_trans + bump_trans
    rot = tt[:, None] * end_rot[None, :] + wob_rot + bump_rot
    lim = float(rot_limit_deg)
    rot = np.deg2rad(np.clip(rot, -lim, lim))
    return np.hstack([trans, rot])


def sample_disorder(K, U, R=1, seed=42):
    """simulation.py:24-63: DISORDER by NumPy permutations.  Sample k = (ky,
    kx) of the K grid belongs to tile (ky // U0) * (K1 // U1) + kx // U1; for
    each of the (K0 // U0) * (K1 // U1) tiles in turn its samples are permuted
    by the generator, every R-th kept and numbered 0 .. prod(U)//R - 1 as
    shot ids.  Returns (shot_ids (-1 where unsampled), temporal order of the
    flat indices shot by shot, y_order, x_order).  A tile whose kept count is
    not prod(U)//R raises ValueError, as the reference's assignment does."""
    if not float(R).is_integer() or R < 1:
        raise TypeError("R must be a positive integer (the reference slices with it)")
    R = int(R)
    K = (int(K[0]), int(K[1]))
    ky, kx = np.meshgrid(np.arange(K[0]), np.arange(K[1]), indexing="ij")
    ky, kx = ky.ravel(), kx.ravel()
    rng = np.random.default_rng(seed)
    ntx = K[1] // U[1]
    ntiles = (K[0] // U[0]) * ntx
    tile = (ky // U[0]).astype(int) * ntx + (kx // U[1]).astype(int)
    shot_ids = -np.ones(K[0] * K[1], dtype=int)
    nshots = int(np.prod(U)) // R
    for t in range(ntiles):
        ids = rng.permutation(np.flatnonzero(tile == t))[::R]
        shot_ids[ids] = np.arange(nshots, dtype=int)
    temporal = np.concatenate([np.flatnonzero(shot_ids == s) for s in range(nshots)])
    y_order, x_order = np.unravel_index(temporal, K)
    return shot_ids, temporal, y_order.astype(int), x_order.astype(int)


def sequential_masks(K, nshots, accel):
    """simulation.py:412-433: line-wise order, partitions 0, accel, 2 accel,
    ... each with all K1 lines; shot s takes samples [s*n, (s+1)*n) with
    n = len(order) // nshots (the remainder is dropped).  Returns (nshots, K0,
    K1, 1) bool.  K0 must be a multiple of accel (the reference's two index
    arrays differ in length otherwise; ValueError here)."""
    K0, K1 = int(K[0]), int(K[1])
    y = np.repeat(np.arange(0, K0, accel), K1)
    x = np.tile(np.arange(K1), K0 // accel)
    if y.shape != x.shape:
        raise ValueError(f"K0 = {K0} is not a multiple of accel = {accel}")
    A = np.zeros((nshots, K0, K1), dtype=bool)
    n = y.shape[0] // nshots
    for s in range(nshots):
        A[s, y[s * n:(s + 1) * n], x[s * n:(s + 1) * n]] = True
    return A[..., None]


def _complex_max(x):
    """NumPy's max of a complex array (lexicographic: largest real part, then
    largest imaginary part among those), as a complex64 scalar; x a complex
    torch tensor (evaluated where it lives)."""
    re = x.real.max()
    im = x.imag[x.real == re].max()
    return np.complex64(complex(float(re), float(im)))


def run_experiment(gt, mps, tile_size, accel=2, order="disorder", iters=200, noise=0, mask=None, low_freq_var=1,
                   continuous=False, out_dir=None, sense_iter=100, pyramid_kwargs=None):
    """The synthetic experiment of simulation.py:250-493 (without its CLI and
    PNG previews).

    gt (N0, N1, N2) and mps (C, N0, N1, N2) centred (NumPy, torch or
    CuPy-like).  Steps, as the reference:

    1. nshots = prod(tile_size) // accel recon shots; corruption states
       N1 // accel with ``continuous``, else nshots;
    2. masks: ``mask`` (shots, K0, K1), used for both the reconstruction and
       the corruption (with the continuous option ignored, as in the
       reference), else DISORDER (order seed 123456) or line-wise
       ``"sequential"``, binned per shot;
    3. motion: generate_motion_parameters_new(states, low_freq_var,
       high_freq_var=20);
    4. corrupted k-space; with ``noise`` > 0 plus complex Gaussian noise of
       std = max(ksp) * 1e-4 * noise (max in NumPy's complex order; the
       noise from NumPy's generator, seed 42);
    5. the uncorrected SENSE image (preconditioner 1 / (sum |S|^2 + 1e-6),
       tol 1e-12, ``sense_iter`` iterations, SigPy's default 100);
    6. pyramid_reconstruction(ksp, mps, A, nshots, n_joint_iters=iters) plus
       ``pyramid_kwargs`` (the reference passes none, so its pyramid
       downsamples the DISORDER way for sequential order too);
    7. err_volume = |corrected - gt|, err_img its 2-norm.

    With ``out_dir``: corrected.nii.gz / uncorrected.nii.gz (magnitudes,
    float64), results.npz (simulation.py:488-493) and the pyramid's loss
    histories.  Returns a dict with corrected, uncorrected, t_estimates,
    err_volume, err_img, the masks, the corruption motion and the k-space."""
    from .io import save_nifti
    from .pyramid import pyramid_reconstruction
    from .sense import sense_recon

    dev = _lib.require_gpu()
    gt_d = as_c64(gt, dev)
    mps_d = as_c64(mps, dev)
    if gt_d.dim() != 3:
        raise AssertionError("gt must be 3-D")
    if mps_d.dim() != 4 or tuple(mps_d.shape[1:]) != tuple(gt_d.shape):
        raise AssertionError("mps must be (coils, *gt.shape)")
    tile = tuple(int(v) for v in tile_size)
    if len(tile) != 2:
        raise AssertionError("tile_size must hold two integers (partitions, lines)")
    shape = tuple(gt_d.shape)
    if shape[0] % tile[0] or shape[1] % tile[1]:
        raise AssertionError("gt's first two axes must be multiples of the tile")
    nshots = tile[0] * tile[1] // accel
    nstates = shape[1] // accel if continuous else nshots
    K = shape[:2]
    if mask is not None:
        A = np.asarray(mask, dtype=bool)[..., None]
        A_corr = A
    elif order == "disorder":
        cfg = DisorderConfig(max_line_number=shape[1] - 1, max_partition_number=shape[0] - 1,
                             pat_lines_to_measure=shape[1] // accel, block_lin=tile[1], block_par=tile[0],
                             seed=123456)
        ordr, _ = make_disorder_order(cfg)
        A = disorder_masks(ordr, nshots, K)
        A_corr = disorder_masks(ordr, nstates, K)
    elif order == "sequential":
        A = sequential_masks(K, nshots, accel)
        A_corr = sequential_masks(K, nstates, accel)
    else:
        raise ValueError("order must be 'disorder' or 'sequential'")
    t_corr = generate_motion_parameters_new(nstates, low_freq_var=low_freq_var, high_freq_var=20.0)[:nstates]
    ksp = generate_corrupted_kspace(gt_d, mps_d, A_corr, t_corr)
    if noise > 0:
        std = _complex_max(ksp) * np.float32(0.0001) * np.float32(noise)
        ksp = ksp + torch.as_tensor(generate_complex_gaussian_noise(tuple(ksp.shape), std=std), device=dev)
    ssq = (mps_d.abs() ** 2).sum(0)
    P = (1.0 / (ssq + 1e-6)).cpu().numpy()
    uncorrected = sense_recon(ksp, mps_d, P=P, tol=1e-12, max_iter=sense_iter).cpu().numpy()
    corrected, t_est = pyramid_reconstruction(ksp, mps_d, A, nshots, n_joint_iters=iters, save_path=out_dir,
                                              **(pyramid_kwargs or {}))
    gt_h = gt_d.cpu().numpy()
    err_volume = np.real(np.abs(corrected - gt_h))
    err_img = float(np.linalg.norm(err_volume))
    res = dict(corrected=corrected, uncorrected=uncorrected, t_estimates=t_est, err_volume=err_volume,
               err_img=err_img, A=A, A_corr=A_corr, t_corr=t_corr, ksp=ksp)
    if out_dir is not None:
        import os
        os.makedirs(out_dir, exist_ok=True)
        save_nifti(os.path.join(str(out_dir), "corrected.nii.gz"), np.abs(corrected).astype(np.float64))
        save_nifti(os.path.join(str(out_dir), "uncorrected.nii.gz"), np.abs(uncorrected).astype(np.float64))
        np.savez(os.path.join(str(out_dir), "results"), corrected=np.abs(corrected), uncorrected=np.abs(uncorrected),
                 t_estimates=t_est, err_volume=err_volume, err_img=err_img)
    return res


CONFIGS = {
    # name: shape, coils, tile (par, lin), accel, recon shots, corruption states, motion kind
    "c1": dict(shape=(64, 64, 64), coils=8, tile=(4, 4), accel=1, nshots=16, nstates=16, motion="uniform"),
    "c2": dict(shape=(128, 128, 96), coils=16, tile=(8, 8), accel=2, nshots=32, nstates=32, motion="walk"),
    "c3": dict(shape=(192, 192, 128), coils=24, tile=(8, 12), accel=2, nshots=48, nstates=48, motion="walk"),
    "c5": dict(shape=(256, 256, 192), coils=32, tile=(16, 16), accel=2, nshots=64, nstates=128, motion="cubic"),
}


def make_config(name, device=None, snr_db=30.0, shape=None, coils=None):
    """Synthetic problem of a benchmark configuration: centred ksp/mps/masks
    (recon shots), true per-shot motion and the ground-truth image.  The
    corruption may use more motion states than the reconstruction
    (simulation.py:373, 404-410)."""
    cfg = dict(CONFIGS[name])
    if shape is not None:
        cfg["shape"] = tuple(shape)
    if coils is not None:
        cfg["coils"] = coils
    device = device or _lib.require_gpu()
    shp = cfg["shape"]
    gt = phantom(shp, seed=0, device=device)
    mps = coil_maps(shp, cfg["coils"], seed=1, device=device)
    A, order = disorder_sampling(shp, cfg["tile"], cfg["accel"], nshots=cfg["nshots"])
    Acorr = disorder_masks(order, cfg["nstates"], tuple(shp[:2]))
    if cfg["motion"] == "uniform":
        tc = motion_uniform(cfg["nstates"])
    elif cfg["motion"] == "walk":
        tc = motion_random_walk(cfg["nstates"])
    else:
        tc = motion_smooth_cubic(cfg["nstates"])
    ksp = generate_corrupted_kspace(gt, mps, Acorr, tc)
    if snr_db is not None:
        ksp = add_noise(ksp, snr_db, A.any(0))
    return dict(gt=gt, mps=mps, ksp=ksp, A=A, A_corr=Acorr, t_corr=tc, config=cfg)


BATCH_SHAPES = ((64, 64, 48), (96, 96, 64), (128, 128, 96))


def batch_specs(n_volumes=24, seed=4, shapes=BATCH_SHAPES, coils=(8, 20), shots=(16, 64)):
    """Seeded descriptions of a set of uneven volumes: shape from ``shapes``,
    coil count uniform in ``coils``, and a DISORDER tile (par, lin) with
    sides in {4, 8, 16} and acceleration in {1, 2, 4} whose shot count
    prod(tile) / accel lies in ``shots``; random-walk motion per shot."""
    rng = np.random.default_rng(seed)
    combos = [((a, b), r) for a in (4, 8, 16) for b in (4, 8, 16) for r in (1, 2, 4)
              if shots[0] <= a * b // r <= shots[1] and (a * b) % r == 0]
    out = []
    for v in range(n_volumes):
        shape = tuple(shapes[int(rng.integers(len(shapes)))])
        tile, accel = combos[int(rng.integers(len(combos)))]
        out.append(dict(shape=shape, coils=int(rng.integers(coils[0], coils[1] + 1)), tile=tile, accel=accel,
                        nshots=tile[0] * tile[1] // accel, seed=int(rng.integers(1 << 30))))
    return out


def make_volume(spec, device=None, snr_db=30.0):
    """Synthetic centred problem of one :func:`batch_specs` entry (motion-
    corrupted with the same per-shot random walk it is reconstructed with)."""
    device = device or _lib.require_gpu()
    shp, sd = spec["shape"], spec["seed"]
    gt = phantom(shp, seed=sd % 1000, device=device)
    mps = coil_maps(shp, spec["coils"], seed=sd % 1000 + 1, device=device)
    A, _ = disorder_sampling(shp, spec["tile"], spec["accel"], nshots=spec["nshots"], seed=sd % 100000)
    tc = motion_random_walk(spec["nshots"], seed=sd % 1000 + 2)
    ksp = generate_corrupted_kspace(gt, mps, A, tc)
    if snr_db is not None:
        ksp = add_noise(ksp, snr_db, A.any(0), seed=sd % 1000 + 3)
    return dict(gt=gt, mps=mps, ksp=ksp, A=A, t_corr=tc, spec=spec)
