"""Grid construction (utils.py:107-137), the random-walk motion generator
(utils.py:5-36) and tensor plumbing helpers."""
from __future__ import annotations

import numpy as np
import torch


class GridDict(dict):
    """dict with the same keys as the reference's kgrid/rkgrid, carrying the
    1-D factors the kernels take: ``r`` = (rx, ry, rz) in full-resolution
    voxel units and ``shape`` = the level shape."""

    r = None
    shape = None


def _axis_vectors(n, ds, v):
    # utils.py:114-120: fftshift(linspace(-n//2, n//2, ds) * v) and
    # fftshift(linspace(-pi, pi, ds))
    r = np.fft.fftshift(np.linspace(-n // 2, n // 2, ds, endpoint=False) * v)
    k = np.fft.fftshift(np.linspace(-np.pi, np.pi, ds, endpoint=False))
    return r, k


def compute_transform_grids_voxel(shape, voxel_size, downsample_res=None, xp=np):
    """utils.py:107-137.  Returns (kgrid, rkgrid) with the reference's keys
    and broadcast shapes.  ``xp`` may be numpy (host arrays) or torch (device
    float64 tensors).  Both dicts also carry ``.r`` (1-D rx, ry, rz, device
    float64 tensors when a GPU is present) and ``.shape`` (level shape)."""
    nx, ny, nz = shape
    vx, vy, vz = voxel_size
    if downsample_res is None:
        downsample_res = shape
    dsx, dsy, dsz = downsample_res
    rx, kx = _axis_vectors(nx, dsx, vx)
    ry, ky = _axis_vectors(ny, dsy, vy)
    rz, kz = _axis_vectors(nz, dsz, vz)
    to = (lambda a: torch.as_tensor(a, dtype=torch.float64, device=_device())) if xp is torch else (lambda a: a)
    kgrid = GridDict(x=to(kx.reshape(-1, 1, 1)), y=to(ky.reshape(1, -1, 1)), z=to(kz.reshape(1, 1, -1)))
    rkgrid = GridDict(
        tan=dict(x=to(rz.reshape(1, 1, -1) * ky.reshape(1, -1, 1)),
                 y=to(rx.reshape(-1, 1, 1) * kz.reshape(1, 1, -1)),
                 z=to(ry.reshape(1, -1, 1) * kx.reshape(-1, 1, 1))),
        sin=dict(x=to(ry.reshape(1, -1, 1) * kz.reshape(1, 1, -1)),
                 y=to(rz.reshape(1, 1, -1) * kx.reshape(-1, 1, 1)),
                 z=to(rx.reshape(-1, 1, 1) * ky.reshape(1, -1, 1))),
    )
    rvec = (rx, ry, rz)
    for g in (kgrid, rkgrid):
        g.r = rvec
        g.shape = (int(dsx), int(dsy), int(dsz))
    return kgrid, rkgrid


def generate_motion_parameters(num_states, low_freq_var=0.1, high_freq_var=5.0, spike_prob=0.02, seed=123456789):
    """utils.py:5-36: a Gaussian random walk per parameter (variance
    ``low_freq_var`` per step) plus sparse spikes (each entry with probability
    ``spike_prob``, Gaussian of variance ``high_freq_var``); the last three
    columns are then read as degrees and converted to radians (no clipping).
    Same NumPy generator and draw order as the reference, so the output is
    identical for a given seed.  Returns (num_states, 6) float64."""
    rng = np.random.default_rng(seed)
    shape = (num_states, 6)
    walk = np.cumsum(rng.normal(0, np.sqrt(low_freq_var), size=shape), axis=0)
    spike_on = rng.choice([0, 1], size=shape, p=[1 - spike_prob, spike_prob])
    spike_amp = rng.normal(0, np.sqrt(high_freq_var), size=shape)
    out = walk + spike_on * spike_amp
    out[:, 3:] *= np.pi / 180
    return out


def _device():
    return torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() else torch.device("cpu")


def _to_numpy(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    if hasattr(a, "get"):  # CuPy-like
        return a.get()
    return np.asarray(a)


def grid_vectors(kgrid, rkgrid):
    """(level shape, (rx, ry, rz) as numpy float64) from grid dicts built by
    this module or by the reference's compute_transform_grids_voxel.

    Reference-built dicts carry only the broadcast planes; r is recovered as
    rk/k at the first non-zero frequency.  k must be the standard grid
    2*pi*fftfreq (Nyquist at -pi); anything else raises ValueError."""
    if getattr(kgrid, "r", None) is not None and getattr(kgrid, "shape", None) is not None:
        return tuple(kgrid.shape), tuple(np.asarray(v, dtype=np.float64) for v in kgrid.r)
    kx = _to_numpy(kgrid["x"]).reshape(-1)
    ky = _to_numpy(kgrid["y"]).reshape(-1)
    kz = _to_numpy(kgrid["z"]).reshape(-1)
    shape = (kx.size, ky.size, kz.size)
    for k in (kx, ky, kz):
        m = k.size
        j = np.arange(m)
        std = 2 * np.pi * (j - m * (j >= m // 2)) / m
        if m % 2 or not np.allclose(k, std, rtol=0, atol=1e-9):
            raise ValueError("kgrid must be fftshift(linspace(-pi, pi, m)) with even m (utils.py:118-120)")
    tz = _to_numpy(rkgrid["tan"]["z"])  # ry * kx  (N0, N1, 1)
    sz = _to_numpy(rkgrid["sin"]["z"])  # rx * ky  (N0, N1, 1)
    tx = _to_numpy(rkgrid["tan"]["x"])  # rz * ky  (1, N1, N2)
    rx = sz[:, 1, 0].real / ky[1]
    ry = tz[1, :, 0].real / kx[1]
    rz = tx[0, 1, :].real / ky[1]
    return shape, (rx.astype(np.float64), ry.astype(np.float64), rz.astype(np.float64))


def as_c64(x, device=None):
    """Contiguous complex64 device tensor view/copy of ``x`` (torch, numpy or
    CuPy-like)."""
    device = device or _device()
    if isinstance(x, torch.Tensor):
        t = x.to(device=device)
    else:
        t = torch.as_tensor(_to_numpy(x), device=device)
    if t.dtype != torch.complex64:
        t = t.to(torch.complex64)
    return t.contiguous()


def as_f64(x, device=None):
    device = device or _device()
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.asarray(_to_numpy(x), dtype=np.float64), device=device).contiguous()
