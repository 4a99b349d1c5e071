"""Image-quality metrics of the experiment scripts on the GPU
(simulation.py:87-97 gradient entropy, simulation.py:476-477 error volume)."""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .utils import _to_numpy, as_c64

NBINS = 256


def error_volume(corrected, gt):
    """err_volume = |corrected - gt| (float32) and err_img = its 2-norm
    (simulation.py:476-477)."""
    dev = _lib.require_gpu()
    a = as_c64(corrected, dev).contiguous()
    b = as_c64(gt, dev).contiguous()
    if a.shape != b.shape:
        raise ValueError(f"shape mismatch {tuple(a.shape)} vs {tuple(b.shape)}")
    out = torch.empty(a.shape, dtype=torch.float32, device=dev)
    part = torch.empty(int(_lib.size("me_metric_partials_size")), dtype=torch.float64, device=dev)
    ss = torch.empty(1, dtype=torch.float64, device=dev)
    _lib.call("me_abs_diff_norm", _lib.ptr(a), _lib.ptr(b), a.numel(), _lib.ptr(out), _lib.ptr(part), _lib.ptr(ss),
              _lib.stream(dev))
    return out, float(np.sqrt(ss.item()))


def _edges(vmax):
    """NumPy's bin edges for histogram(..., bins=256, range=(0, vmax + 1e-8))
    of float32 data: the range end is a float32 (NumPy 2 promotion), the
    edges float32 from linspace."""
    hi = np.float32(vmax) + 1e-8
    lo = 0
    if lo == hi:  # numpy widens an empty range
        lo, hi = lo - 0.5, hi + 0.5
    bt = np.result_type(lo, hi, np.float32)
    return np.linspace(lo, hi, NBINS + 1, endpoint=True, dtype=bt).astype(np.float32)


def gradient_entropy(vol):
    """simulation.py:87-97 for a float32 volume (N0, N1, N2), e.g. |x| of a
    complex64 reconstruction; lower is sharper."""
    dev = _lib.require_gpu()
    if isinstance(vol, torch.Tensor):
        v = vol.detach().to(device=dev, dtype=torch.float32).contiguous()
    else:
        v = torch.as_tensor(np.ascontiguousarray(_to_numpy(vol), dtype=np.float32), device=dev)
    if v.dim() != 3:
        raise ValueError("gradient_entropy expects a 3-D volume")
    n = _lib.int3(v.shape)
    s = _lib.stream(dev)
    bits = torch.empty(1, dtype=torch.int32, device=dev)
    _lib.call("me_gradient_max", _lib.ptr(v), n, _lib.ptr(bits), s)
    vmax = float(bits.view(torch.float32).item())
    edges = torch.as_tensor(_edges(vmax), device=dev)
    counts = torch.empty(NBINS, dtype=torch.int64, device=dev)
    ent = torch.empty(1, dtype=torch.float64, device=dev)
    _lib.call("me_gradient_entropy", _lib.ptr(v), n, _lib.ptr(edges), _lib.ptr(counts), _lib.ptr(ent), s)
    gradient_entropy.counts = counts
    gradient_entropy.max = vmax
    return float(ent.item())
