"""Synthetic ESPIRiT inputs shared by the oracle pins and the GPU parity
tests: a smooth object seen through smooth, known, normalised coil maps."""
import numpy as np


def espirit_data(imsize=(24, 20, 16), nc=4, cal=(12, 12, 10), seed=80):
    """k-space of a smooth object seen through smooth known maps; the
    centred calibration region of size ``cal``."""
    rng = np.random.default_rng(seed)
    g = np.meshgrid(*[np.linspace(-1, 1, n) for n in imsize], indexing="ij")
    obj = np.exp(-(g[0] ** 2 + g[1] ** 2 + g[2] ** 2) / 0.5) * (1 + 0.3 * g[0])
    maps = []
    for c in range(nc):
        a = 2 * np.pi * c / nc
        maps.append(np.exp(-((g[0] - 0.8 * np.cos(a)) ** 2 + (g[1] - 0.8 * np.sin(a)) ** 2) / 1.5)
                    * np.exp(1j * (0.4 * g[2] + 0.3 * c)))
    maps = np.stack(maps)
    maps /= np.sqrt(np.sum(np.abs(maps) ** 2, axis=0))
    ksp = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(maps * obj, axes=(1, 2, 3)), axes=(1, 2, 3)), axes=(1, 2, 3))
    ksp += 1e-6 * np.abs(ksp).max() * (rng.standard_normal(ksp.shape) + 1j * rng.standard_normal(ksp.shape))
    sl = tuple(slice((n - m) // 2, (n - m) // 2 + m) for n, m in zip(imsize, cal))
    calib = np.transpose(ksp[(slice(None),) + sl], (1, 2, 3, 0))
    return calib, maps, obj
