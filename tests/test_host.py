"""CPU checks of host-side helpers (no device needed)."""
import numpy as np
import pytest
import torch


def test_aggregate_sampling_masks_counts_lines(me):
    # 6 shots of a 4x2 grid, shot s samples line (s % 4, s // 4); bins of 3
    A = np.zeros((6, 4, 2, 1), dtype=bool)
    for s in range(6):
        A[s, s % 4, s // 4, 0] = True
    agg = me.aggregate_sampling_masks(A, 3)
    assert agg.shape == (2, 4, 2, 1)
    assert agg[0, :3, 0, 0].tolist() == [1, 1, 1] and agg[0].sum() == 3
    assert agg[1, 3, 0, 0] == 1 and agg[1, 0, 1, 0] == 1 and agg[1, 1, 1, 0] == 1
    At = me.aggregate_sampling_masks(torch.as_tensor(A), 3)
    assert np.array_equal(At.numpy(), agg)


def test_aggregate_transforms_sums_bins(me):
    T = np.arange(4 * 6, dtype=float).reshape(4, 6)
    agg = me.aggregate_transforms(T, 2)
    assert agg.shape == (2, 6)
    assert np.array_equal(agg[0], T[0] + T[1]) and np.array_equal(agg[1], T[2] + T[3])
    assert np.array_equal(me.aggregate_transforms(torch.as_tensor(T), 2).numpy(), agg)
    with pytest.raises(ValueError):
        me.aggregate_transforms(np.zeros((5, 6)), 2)
