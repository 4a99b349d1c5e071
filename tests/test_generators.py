"""The reference's seeded NumPy generators, bit-exact against vectors the
reference functions wrote (tests/golden/make_generator_vectors.py):
generate_motion_parameters_new (simulation.py:104-174),
generate_motion_parameters (utils.py:5-36), generate_complex_gaussian_noise
(simulation.py:99-102, NumPy generator) and sample_disorder
(simulation.py:24-63); plus the sequential masks of simulation.py:412-433 by
their definition."""
import os

import numpy as np
import pytest

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "motion_generators.npz"))


def _cases(prefix):
    return sorted({k.split("_")[1] for k in GOLDEN.files if k.startswith(prefix + "_") and k.endswith("_args")})


@pytest.mark.parametrize("i", _cases("new"))
def test_generate_motion_parameters_new_matches_reference(me, i):
    n, lf, hf, sp, sd, lim = GOLDEN[f"new_{i}_args"]
    got = me.simulation.generate_motion_parameters_new(int(n), low_freq_var=lf, high_freq_var=hf, spike_prob=sp,
                                                       seed=int(sd), rot_limit_deg=lim)
    assert np.array_equal(got, GOLDEN[f"new_{i}"])


def test_generate_motion_parameters_new_limits(me):
    t = me.simulation.generate_motion_parameters_new(300, low_freq_var=400, high_freq_var=20.0, rot_limit_deg=2.5)
    assert t.shape == (300, 6)
    assert np.degrees(np.abs(t[:, 3:])).max() <= 2.5 + 1e-12  # rotations clipped, then radians
    assert np.abs(t[0, :3]).max() < 1e-3 * np.abs(t[:, :3]).max()  # drift and wobble start at zero


@pytest.mark.parametrize("i", _cases("old"))
def test_generate_motion_parameters_matches_reference(me, i):
    n, lf, hf, sp, sd = GOLDEN[f"old_{i}_args"]
    got = me.utils.generate_motion_parameters(int(n), low_freq_var=lf, high_freq_var=hf, spike_prob=sp, seed=int(sd))
    assert np.array_equal(got, GOLDEN[f"old_{i}"])


def test_generate_complex_gaussian_noise_matches_reference(me):
    sim = me.simulation
    assert np.array_equal(sim.generate_complex_gaussian_noise((3, 5, 4, 2), std=0.25, seed=42), GOLDEN["noise_0"])
    got = sim.generate_complex_gaussian_noise((2, 7, 3), std=np.complex64(0.5 - 0.125j), seed=9)
    assert got.dtype == np.complex64
    assert np.array_equal(got, GOLDEN["noise_1"])


@pytest.mark.parametrize("i", _cases("dis"))
def test_sample_disorder_matches_reference(me, i):
    k0, k1, u0, u1, r, sd = (int(v) for v in GOLDEN[f"dis_{i}_args"])
    got = me.simulation.sample_disorder((k0, k1), (u0, u1), R=r, seed=sd)
    for arr, name in zip(got, ("shot_ids", "temporal", "y", "x")):
        assert np.array_equal(arr, GOLDEN[f"dis_{i}_{name}"]), name


def test_sample_disorder_invariants(me):
    shot_ids, temporal, y, x = me.simulation.sample_disorder((16, 12), (4, 4), R=2, seed=5)
    # every tile gives each of the 8 shots exactly one sample
    tiles = (np.arange(16)[:, None] // 4) * 3 + np.arange(12)[None, :] // 4
    ids = shot_ids.reshape(16, 12)
    for t in range(12):
        assert sorted(ids[tiles == t][ids[tiles == t] >= 0].tolist()) == list(range(8))
    assert np.array_equal(np.ravel_multi_index((y, x), (16, 12)), temporal)
    with pytest.raises(TypeError):
        me.simulation.sample_disorder((8, 8), (4, 4), R=0.5)


def test_sequential_masks(me):
    A = me.simulation.sequential_masks((8, 6), 4, 2)
    assert A.shape == (4, 8, 6, 1)
    # partitions 0, 2, 4, 6 with all lines, 24 samples, 6 per shot in order
    assert A.sum() == 24 and np.array_equal(A.sum(axis=(1, 2, 3)), [6, 6, 6, 6])
    assert A[0, 0, :, 0].all() and A[1, 2, :, 0].all() and A[3, 6, :, 0].all()
    assert not A[:, 1::2].any()
    with pytest.raises(ValueError):
        me.simulation.sequential_masks((9, 6), 4, 2)
