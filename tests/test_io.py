"""CPU checks of the on-disk formats (recon.py:82-165, simulation.py:476-493):
NIfTI-1 files against the format's own definition (nibabel is absent here,
so byte-level parity with nib.save is unpinned; the fields it writes for
Nifti1Image(data, eye(4)) are checked one by one), the preprocessed-array
loader, the shot masks from scanner counters against the oracle, and the
centred resize against SigPy's rule as restated in the oracle."""
import gzip
import os
import struct

import numpy as np
import pytest
import torch

from oracle import motion_oracle as O


@pytest.mark.parametrize("name", ["v.nii", "v.nii.gz"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.int16])
def test_nifti_round_trip_and_header(me, tmp_path, name, dtype):
    rng = np.random.default_rng(1)
    data = (rng.standard_normal((7, 5, 3)) * 100).astype(dtype)
    path = tmp_path / name
    me.io.save_nifti(str(path), data)
    raw = (gzip.open if name.endswith(".gz") else open)(path, "rb").read()
    assert len(raw) == 352 + data.nbytes
    assert struct.unpack("<i", raw[:4])[0] == 348
    assert raw[344:348] == b"n+1\0"
    assert struct.unpack("<8h", raw[40:56]) == (3, 7, 5, 3, 1, 1, 1, 1)
    code = {np.float64: 64, np.float32: 16, np.int16: 4}[dtype]
    assert struct.unpack("<hh", raw[70:74]) == (code, np.dtype(dtype).itemsize * 8)
    assert struct.unpack("<f", raw[108:112])[0] == 352.0
    assert struct.unpack("<hh", raw[252:256]) == (0, 2)  # qform_code, sform_code
    assert np.array_equal(np.array(struct.unpack("<12f", raw[280:328])).reshape(3, 4), np.eye(4)[:3])
    assert raw[348:352] == b"\0\0\0\0"
    # voxels in Fortran order (x fastest)
    assert np.array_equal(np.frombuffer(raw[352:], dtype=np.dtype(dtype).newbyteorder("<")),
                          data.ravel(order="F"))
    back, hdr = me.io.load_nifti(str(path))
    assert back.dtype == data.dtype and np.array_equal(back, data)
    assert hdr["sform_code"] == 2 and hdr["magic"] == b"n+1\0"


def test_nifti_affine_and_zooms(me, tmp_path):
    aff = np.diag([2.0, 1.5, 3.0, 1.0])
    aff[:3, 3] = [10, -4, 1]
    me.io.save_nifti(str(tmp_path / "a.nii"), np.zeros((2, 2, 2)), affine=aff)
    _, hdr = me.io.load_nifti(str(tmp_path / "a.nii"))
    assert np.allclose(hdr["srow"], aff[:3]) and np.allclose(hdr["pixdim"][1:4], [2.0, 1.5, 3.0])


def test_load_preprocessed(me, tmp_path):
    k = np.ones((2, 3, 4, 5), np.complex64)
    np.save(tmp_path / "ksp.npy", k)
    np.save(tmp_path / "noise.npy", np.zeros((2, 9), np.complex64))
    d = me.io.load_preprocessed(str(tmp_path))
    assert np.array_equal(d["ksp"], k) and d["ref"] is None and d["noise"].shape == (2, 9)


@pytest.mark.parametrize("n_shots", [1, 3, 7])
def test_shot_masks_from_counters_match_oracle(me, n_shots):
    rng = np.random.default_rng(n_shots)
    n_par, n_lin = 6, 10
    samples = rng.permutation(n_par * n_lin)[:45]
    parts, lines = samples // n_lin, samples % n_lin
    got = me.io.shot_masks_from_counters(lines, parts, n_shots, n_par, n_lin)
    ref = O.shot_masks_from_counters(lines, parts, n_shots, n_par, n_lin)
    assert got.dtype == bool and np.array_equal(got, ref)


@pytest.mark.parametrize("shape_in,shape_out", [((4, 30, 26, 31), (4, 24, 24, 24)), ((3, 20, 22, 19), (3, 24, 24, 24)),
                                                ((2, 9, 30, 24), (2, 24, 24, 24))])
def test_centred_resize_matches_sigpy_rule(me, shape_in, shape_out):
    rng = np.random.default_rng(2)
    x = (rng.standard_normal(shape_in) + 1j * rng.standard_normal(shape_in)).astype(np.complex64)
    got = me.io._centred_resize(torch.as_tensor(x), shape_out).numpy()
    assert np.array_equal(got, O.sp_resize(x, shape_out))
