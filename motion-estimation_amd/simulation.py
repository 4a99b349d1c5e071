"""Synthetic inputs: DISORDER sampling (simulation.py:192-245, 383-410), the
motion-corrupted forward model (simulation.py:65-85), and the seeded
phantom / coil / motion / noise generators the benchmark configurations
are defined with.

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
