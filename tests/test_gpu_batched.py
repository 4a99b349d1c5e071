"""Batched multi-volume reconstruction and the line-pass autotuner on the GPU.

* Every lines-per-workgroup variant of the plain line pass gives bitwise
  the same rigid transform (each line's arithmetic does not depend on how
  many lines a workgroup holds).
* The autotuner picks a built variant and records it in the table.
* Volumes of different matrix size, coil count and shot count advanced by
  BatchedJointRecon (one eager outer iteration, then replays of the recorded
  HIP graph over concurrent streams) end bitwise where the same volumes
  advanced one at a time by JointRecon end."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _geom(me, shape):
    from motion_estimation_amd.utils import compute_transform_grids_voxel
    kg, rkg = compute_transform_grids_voxel(shape, (1, 1, 1))
    return me.transform.Geometry.from_grids(kg, rkg, DEV)


@pytest.mark.parametrize("shape", [(48, 32, 24), (64, 64, 48)])
def test_line_variants_bitwise_equal(me, shape):
    g = _geom(me, shape)
    rng = np.random.default_rng(1)
    B = 3
    x = torch.as_tensor((rng.standard_normal((B, *shape)) + 1j * rng.standard_normal((B, *shape)))
                        .astype(np.complex64), device=DEV)
    theta = torch.as_tensor(rng.uniform(-0.05, 0.05, (B, 6)) * [40, 40, 40, 1, 1, 1], device=DEV)
    outs = {}
    try:
        for lines in (16, 8, 32):
            for plane in (False, True):
                st = g.st_plane if plane else g.st_std
                for ax in range(3):
                    me._lib.call("me_line_tuning_set", me._lib.int3(shape), me._lib.ll3(st), ax, lines)
            outs[lines] = [me.transform.rigid_batch(x, theta, g, mode=m, plane=p).cpu()
                           for m in (0, 1) for p in (False, True)]
    finally:
        me._lib.load().me_line_tuning_clear()
    for lines in (8, 32):
        for a, b in zip(outs[lines], outs[16]):
            assert torch.equal(a, b), lines


def test_line_tuning_rejects_unbuilt_variant(me):
    with pytest.raises(RuntimeError):
        me._lib.call("me_line_tuning_set", me._lib.int3((64, 64, 48)), me._lib.ll3((64, 1, 4096)), 0, 12)
    assert me._lib.size("me_line_tuning_get", me._lib.int3((64, 64, 48)), me._lib.ll3((64, 1, 4096)), 0) == 16


def test_tuner_picks_built_variant(me):
    shape = (64, 64, 48)
    n0, n1, n2 = shape
    buf = torch.zeros((4, n0 * n1 * n2), dtype=torch.complex64, device=DEV)
    chosen = (C.c_int * 3)()
    times = (C.c_float * 9)()
    st = me._lib.ll3((n1, 1, n0 * n1))
    try:
        me._lib.call("me_tune_line_passes", me._lib.ptr(buf), me._lib.ptr(buf), n0 * n1 * n2, 4,
                     me._lib.int3(shape), st, 3, chosen, times, me._lib.stream(DEV))
        for ax in range(3):
            assert chosen[ax] in (8, 16, 32)
            assert me._lib.size("me_line_tuning_get", me._lib.int3(shape), st, ax) == chosen[ax]
            t = [times[3 * ax + c] for c in range(3)]
            assert t[1] > 0 and min(v for v in t if v > 0) == times[3 * ax + (8, 16, 32).index(chosen[ax])]
    finally:
        me._lib.load().me_line_tuning_clear()


def _volumes(me, specs):
    from motion_estimation_amd.utils import compute_transform_grids_voxel
    out = []
    for sp in specs:
        d = me.simulation.make_volume(sp, device=DEV)
        kg, rkg = compute_transform_grids_voxel(sp["shape"], (1, 1, 1))
        out.append(dict(ksp=d["ksp"], mps=d["mps"], shot_mask=d["A"], kgrid=kg, rkgrid=rkg))
    return out


def test_batched_graph_matches_one_at_a_time(me):
    specs = [dict(shape=(32, 32, 24), coils=5, tile=(4, 4), accel=1, nshots=16, seed=11),
             dict(shape=(48, 48, 32), coils=9, tile=(8, 4), accel=2, nshots=16, seed=12),
             dict(shape=(64, 64, 48), coils=12, tile=(8, 8), accel=2, nshots=32, seed=13),
             dict(shape=(32, 32, 24), coils=8, tile=(4, 8), accel=1, nshots=32, seed=14)]
    vols = _volumes(me, specs)
    n_iter = 4
    bj = me.BatchedJointRecon(vols, nstreams=3)
    try:
        # tune first: the process-wide tables then serve the eager reference too
        picks = bj.tune(reps=2)
        assert set(picks) == {sp["shape"] for sp in specs}
        assert len(bj.encoder_tuning) == len({(sp["shape"], sp["coils"]) for sp in specs})
        for enc in bj.encoder_tuning.values():
            for variant, ms in enc.values():
                assert variant in (0, 1, 2, 3) and min(ms) > 0
        ref = []
        for v in vols:
            jr = me.JointRecon(v["ksp"], v["mps"], v["shot_mask"], v["kgrid"], v["rkgrid"], max_joint_iter=n_iter,
                               compute_objective=False)
            for _ in range(n_iter):
                jr.step()
            ref.append((jr.image().cpu(), jr.t.cpu()))
        imgs, ts = bj.run(n_iter)
        torch.cuda.synchronize()
        assert bj.steps_done == n_iter and bj._exec is not None
        for (ri, rt), gi, gt in zip(ref, imgs, ts):
            assert torch.equal(gi.cpu(), ri)
            assert torch.equal(gt.cpu(), rt)
            assert float(rt.abs().max()) > 0  # the motion estimate moved
    finally:
        bj.close()
        me._lib.load().me_line_tuning_clear()
        me._lib.load().me_encode_tuning_clear()
