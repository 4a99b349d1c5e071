"""DISORDER order (simulation.py:201-245): the library's C implementation
against the oracle (stdlib random.Random, as the reference) -- bit-exact --
and both against orders written by the reference's own make_disorder_order
(tests/golden/disorder_orders.npz, made by tests/golden/make_generator_vectors.py)."""
import os

import numpy as np
import pytest

from oracle import motion_oracle as O

GOLDEN = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "disorder_orders.npz"))
CASES = sorted({int(k.split("_")[1]) for k in GOLDEN.files})


def _cfg(mod, args):
    k0, k1, u0, u1, r, sd = (int(v) for v in args)
    return mod.DisorderConfig(max_line_number=k1 - 1, max_partition_number=k0 - 1, pat_lines_to_measure=k1 // r,
                              block_lin=u1, block_par=u0, seed=sd)


@pytest.mark.parametrize("case", CASES)
def test_oracle_order_equals_reference_vectors(case):
    order, LP = O.make_disorder_order(_cfg(O, GOLDEN[f"case_{case}_args"]))
    assert list(LP) == GOLDEN[f"case_{case}_LP"].tolist()
    assert np.array_equal(np.array(order, dtype=np.int64), GOLDEN[f"case_{case}_order"].astype(np.int64))


@pytest.mark.parametrize("case", CASES)
def test_library_order_equals_reference_vectors(me, case):
    order, LP = me.simulation.make_disorder_order(_cfg(me.simulation, GOLDEN[f"case_{case}_args"]))
    assert list(LP) == GOLDEN[f"case_{case}_LP"].tolist()
    assert np.array_equal(np.array(order, dtype=np.int64), GOLDEN[f"case_{case}_order"].astype(np.int64))


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


def test_disorder_masks_drop_leftover_samples(me):
    """simulation.py:396-402: len(order) // nshots samples per shot in order,
    mask[s, partition, line]; the remainder is dropped (10 samples, 3 shots:
    3 each, the tenth unused)."""
    order = [(0, 0), (1, 0), (2, 1), (3, 1), (0, 2), (1, 3), (2, 2), (3, 3), (1, 1), (2, 3)]
    expect = np.zeros((3, 4, 4), dtype=bool)
    for s in range(3):
        for line, part in order[3 * s:3 * s + 3]:
            expect[s, part, line] = True
    for mod in (O, me.simulation):
        got = mod.disorder_masks(order, 3, (4, 4))
        assert got.shape == (3, 4, 4, 1)
        assert np.array_equal(got[..., 0], expect)
        assert not got[:, 3, 2].any()  # the tenth sample (line 2, partition 3)
