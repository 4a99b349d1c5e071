"""DISORDER order (simulation.py:201-245): the library's C implementation
against the oracle (stdlib random.Random, as the reference) -- bit-exact."""
import numpy as np
import pytest

from oracle import motion_oracle as O


@pytest.mark.parametrize("K,tile,accel,seed", [
    ((64, 64), (4, 4), 1, 123456), ((128, 128), (8, 8), 2, 123456), ((256, 256), (16, 16), 2, 123456),
    ((48, 64), (4, 8), 2, 7), ((30, 50), (4, 8), 3, 2**40 + 5), ((192, 192), (8, 12), 2, 0)])
def test_disorder_order_bit_exact(me, K, tile, accel, seed):
    kw = dict(max_line_number=K[1] - 1, max_partition_number=K[0] - 1, pat_lines_to_measure=K[1] // accel,
              block_lin=tile[1], block_par=tile[0], seed=seed)
    got, gLP = me.simulation.make_disorder_order(me.simulation.DisorderConfig(**kw))
    ref, rLP = O.make_disorder_order(O.DisorderConfig(**kw))
    assert gLP == rLP
    assert got == ref


def test_disorder_masks_match_oracle(me):
    order, _ = O.make_disorder_order(O.DisorderConfig(255, 255, 128, 16, 16))
    a = me.simulation.disorder_masks(order, 64, (256, 256))
    b = O.disorder_masks(order, 64, (256, 256))
    assert np.array_equal(a, b)
    assert a.shape == (64, 256, 256, 1) and int(a.sum()) == 32768


def test_disorder_rejects_bad_tiles(me):
    with pytest.raises(RuntimeError):
        me.simulation.make_disorder_order(me.simulation.DisorderConfig(15, 15, 16, 0, 4))
