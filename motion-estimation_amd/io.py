"""On-disk formats either side of the reconstruction (recon.py:82-165,
extract_from_dat.py:6-60, simulation.py:476-493).

* ``load_preprocessed``: the ksp / ref / noise arrays that the reference's
  TWIX extraction writes as .npy (reading the Siemens .dat itself needs
  ``twixtools``, a third-party reader this package does not include).
* ``shot_masks_from_counters``: DISORDER shot masks from the scanner's
  line/partition counters (recon.py:158-165).
* ``prepare_real_data``: the calibration chain of recon.py:130-150 on the
  GPU (noise whitening, ESPIRiT from the reference scan cropped to 24^3,
  coil compression).
* ``reconstruct_real_data``: the rest of recon.py:130-172 from those arrays
  (SENSE image, pyramid motion correction over the counter masks, the pre/
  and recon/ outputs).
* ``save_nifti`` / ``load_nifti``: NIfTI-1 single files (.nii or .nii.gz)
  without nibabel, as nib.save(Nifti1Image(data, np.eye(4))) is used at
  simulation.py:480-481.
* ``save_results``: the npz of simulation.py:488-493.
"""
from __future__ import annotations

import gzip
import os
import struct

import numpy as np
import torch

from . import _lib
from .calibration import coil_compression, noise_whiten
from .espirit import rx_espirit_3d
from .utils import _to_numpy, as_c64

# ------------------------------------------------------------------ TWIX output
def load_preprocessed(directory):
    """ksp.npy, ref.npy and noise.npy as written by extract_from_dat.py:26-56
    (ksp and ref transposed to (coils, par, lin, col)); absent files are None."""
    out = {}
    for name in ("ksp", "ref", "noise"):
        path = os.path.join(directory, f"{name}.npy")
        out[name] = np.load(path) if os.path.exists(path) else None
    return out


def shot_masks_from_counters(lines, partitions, n_shots, n_par, n_lin):
    """recon.py:158-165: shot s owns samples [s*sps, (s+1)*sps) of the
    acquisition order (sps = len(lines) // n_shots; the remainder is
    dropped), marked at (partition, line).  Returns (n_shots, n_par, n_lin, 1)
    bool."""
    lines = np.asarray(lines, dtype=np.int64)
    partitions = np.asarray(partitions, dtype=np.int64)
    if lines.shape != partitions.shape:
        raise ValueError("lines and partitions differ in length")
    mask = np.zeros((n_shots, n_par, n_lin), dtype=bool)
    sps = lines.size // n_shots
    if sps > 0:
        shot = np.repeat(np.arange(n_shots), sps)
        mask[shot, partitions[:n_shots * sps], lines[:n_shots * sps]] = True
    return mask[..., None]


def _centred_resize(x, shape):
    """SigPy resize semantics on the device: centred crop and/or zero pad per
    axis (ishift = max(i//2 - o//2, 0), oshift = max(o//2 - i//2, 0))."""
    out = torch.zeros(shape, dtype=x.dtype, device=x.device)
    src, dst = [], []
    for i, o in zip(x.shape, shape):
        ish, osh = max(i // 2 - o // 2, 0), max(o // 2 - i // 2, 0)
        n = min(i - ish, o - osh)
        src.append(slice(ish, ish + n))
        dst.append(slice(osh, osh + n))
    out[tuple(dst)] = x[tuple(src)]
    return out


def prepare_real_data(ksp, ref, noise, lines=None, partitions=None, n_shots=None, calib_size=24,
                      kernel_size=(5, 5, 5), eig_thresh=0.02, mask_thresh=0.99, perc=0.95, keep_intermediate=False):
    """recon.py:130-165 on the GPU: whiten ksp with the noise covariance, ESPIRiT
    maps from the reference scan cropped/padded to calib_size^3 (not whitened,
    as in the reference), coil compression of maps and whitened k-space, and
    the shot masks from the counters when given.  ksp/ref: (coils, par, lin,
    col); noise: coils first.  Returns dict(ksp, mps, shot_mask, W), with
    ``keep_intermediate`` also white_ksp and the uncompressed maps mps_full."""
    dev = _lib.require_gpu()
    k = as_c64(ksp, dev)
    C = k.shape[0]
    white, W = noise_whiten(k, as_c64(noise, dev).reshape(C, -1))
    del k
    r = _centred_resize(as_c64(ref, dev), (C, calib_size, calib_size, calib_size))
    mps = rx_espirit_3d(r.permute(1, 2, 3, 0).contiguous(), tuple(white.shape[1:]), kernel_size=kernel_size,
                        eig_thresh=eig_thresh, mask_thresh=mask_thresh)
    mps = mps.permute(3, 0, 1, 2).contiguous()
    cc_mps, cc_ksp = coil_compression(mps, white, perc=perc)
    out = dict(ksp=cc_ksp, mps=cc_mps, W=W, shot_mask=None)
    if keep_intermediate:
        out.update(white_ksp=white, mps_full=mps)
    if lines is not None:
        out["shot_mask"] = shot_masks_from_counters(lines, partitions, n_shots, white.shape[1], white.shape[2])
    return out


def reconstruct_real_data(ksp, ref, noise, lines, partitions, n_shots, out_dir=None, sampling="disorder", iters=1000,
                          calib_size=24, sense_iter=100, pyramid_kwargs=None):
    """recon.py:130-172 from the arrays the TWIX extraction yields (ksp and
    ref as (coils, par, lin, col), noise coils first, the cLin / cPar
    counters of the imaging readouts in acquisition order):

    1. :func:`prepare_real_data`: whitening, ESPIRiT maps (kernel 5^3,
       threshold 0.02, mask 0.99) from the reference scan resized to
       calib_size^3, coil compression to 95% and the shot masks;
    2. the SENSE image of the compressed data, SigPy's defaults (no
       preconditioner, ``sense_iter`` = 100 iterations, rss(y) > 0 weights);
    3. pyramid_reconstruction(ksp, mps, masks, n_shots, n_joint_iters=iters,
       sampling=sampling) plus ``pyramid_kwargs`` (the reference passes none).

    With ``out_dir``: pre/ holds ksp, ref, noise, white_ksp, mps, cc_mps,
    cc_ksp (.npy) and recon/ holds sense.npy, joint_recon.npy,
    estimated_transforms.npy, joint_recon.nii.gz (magnitude, float64) and
    the pyramid's loss histories, as the reference writes them.  Returns a
    dict with the prepared arrays, sense, recon and t_est."""
    from .pyramid import pyramid_reconstruction
    from .sense import sense_recon

    prep = prepare_real_data(ksp, ref, noise, lines=lines, partitions=partitions, n_shots=n_shots,
                             calib_size=calib_size, keep_intermediate=True)
    sense = sense_recon(prep["ksp"], prep["mps"], max_iter=sense_iter).cpu().numpy()
    rec_dir = None if out_dir is None else os.path.join(str(out_dir), "recon")
    recon, t_est = pyramid_reconstruction(prep["ksp"], prep["mps"], prep["shot_mask"], n_shots, n_joint_iters=iters,
                                          save_path=rec_dir, sampling=sampling, **(pyramid_kwargs or {}))
    if out_dir is not None:
        pre = os.path.join(str(out_dir), "pre")
        os.makedirs(pre, exist_ok=True)
        os.makedirs(rec_dir, exist_ok=True)
        host = lambda x: np.asarray(_to_numpy(x))  # noqa: E731
        for name, arr in (("ksp", ksp), ("ref", ref), ("noise", noise), ("white_ksp", prep["white_ksp"]),
                          ("mps", prep["mps_full"]), ("cc_mps", prep["mps"]), ("cc_ksp", prep["ksp"])):
            np.save(os.path.join(pre, f"{name}.npy"), host(arr))
        np.save(os.path.join(rec_dir, "sense.npy"), sense)
        np.save(os.path.join(rec_dir, "joint_recon.npy"), recon)
        np.save(os.path.join(rec_dir, "estimated_transforms.npy"), t_est)
        save_nifti(os.path.join(rec_dir, "joint_recon.nii.gz"), np.abs(recon).astype(np.float64))
    return dict(prep, sense=sense, recon=recon, t_est=t_est)


# ------------------------------------------------------------------ NIfTI-1
_HDR = struct.Struct("<i10s18sihcb8hfffhhhh8ffffhcbffffii80s24shhffffff4f4f4f16s4s")
_DTYPES = {np.dtype(np.uint8): 2, np.dtype(np.int16): 4, np.dtype(np.int32): 8, np.dtype(np.float32): 16,
           np.dtype(np.float64): 64, np.dtype(np.complex64): 32}


def _nifti_header(shape, dtype, affine):
    dt = _DTYPES[np.dtype(dtype)]
    dims = [len(shape)] + list(shape) + [1] * (7 - len(shape))
    a = np.asarray(affine, dtype=np.float64)
    zooms = np.sqrt((a[:3, :3] ** 2).sum(axis=0))
    pixdim = [1.0, *zooms.tolist(), 0.0, 0.0, 0.0, 0.0]
    return _HDR.pack(
        348, b"", b"", 0, 0, b"r", 0, *dims,
        0.0, 0.0, 0.0, 0, dt, np.dtype(dtype).itemsize * 8, 0,
        *pixdim, 352.0, 1.0, 0.0, 0, b"\0", 0,
        0.0, 0.0, 0.0, 0.0, 0, 0, b"", b"",
        0, 2,  # qform unknown, sform aligned (the affine)
        0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
        *a[0].tolist(), *a[1].tolist(), *a[2].tolist(),
        b"", b"n+1\0")


def save_nifti(path, data, affine=None):
    """Single-file NIfTI-1 (gzip when the name ends in .gz): 348-byte header,
    a zero extension flag, then the voxels in Fortran order at offset 352,
    the affine in the sform (code 2) as nib.Nifti1Image(data, affine) does."""
    arr = np.asarray(_to_numpy(data))
    if arr.dtype not in _DTYPES:
        raise TypeError(f"unsupported NIfTI data type {arr.dtype}")
    if not 1 <= arr.ndim <= 7:
        raise ValueError("NIfTI holds 1 to 7 dimensions")
    affine = np.eye(4) if affine is None else affine
    body = arr.astype(arr.dtype.newbyteorder("<")).tobytes(order="F")
    blob = _nifti_header(arr.shape, arr.dtype, affine) + b"\0\0\0\0" + body
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "wb") as f:
        f.write(blob)


def load_nifti(path):
    """Inverse of save_nifti for single-file NIfTI-1: (data, header dict)."""
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rb") as f:
        blob = f.read()
    fields = _HDR.unpack(blob[:348])
    if fields[0] != 348 or fields[-1][:3] != b"n+1":
        raise ValueError("not a single-file NIfTI-1")
    ndim = fields[7]
    shape = tuple(fields[8:8 + ndim])
    code = fields[19]
    dtype = {v: k for k, v in _DTYPES.items()}[code]
    off = int(fields[30])
    data = np.frombuffer(blob, dtype=dtype.newbyteorder("<"), offset=off, count=int(np.prod(shape)))
    hdr = dict(dim=fields[7:15], datatype=code, bitpix=fields[20], pixdim=fields[22:30], vox_offset=fields[30],
               scl_slope=fields[31], scl_inter=fields[32], qform_code=fields[44], sform_code=fields[45],
               srow=np.array(fields[52:64]).reshape(3, 4), magic=fields[-1])
    return data.reshape(shape, order="F"), hdr


# ------------------------------------------------------------------ results
def save_results(path, corrected, uncorrected, t_estimates, gt):
    """simulation.py:476-493: err_volume = |corrected - gt| and err_img its
    2-norm (on the GPU), then corrected/uncorrected magnitudes, the motion
    estimates and both errors in one npz."""
    from .metrics import error_volume
    err_volume, err_img = error_volume(corrected, gt)
    mag = lambda x: np.abs(np.asarray(_to_numpy(x)))  # noqa: E731
    np.savez(path, corrected=mag(corrected), uncorrected=mag(uncorrected),
             t_estimates=np.asarray(_to_numpy(t_estimates)), err_volume=err_volume.cpu().numpy(), err_img=err_img)
    return err_img
