"""Batching: the 24 seeded volumes of uneven size (64 x 64 x 48,
96 x 96 x 64, 128 x 128 x 96), coil count (8-20), shot count (16-64) and
DISORDER tile (simulation.batch_specs(24, seed=4), random-walk motion, 30 dB
noise), each advanced by outer iterations of 3 CG steps and one LM update.

* After one eager outer iteration, the cheapest volume (fewest voxels x
  coils x shots) against the FP64 oracle's JointRecon: image 1e-4, motion
  1e-3, damping identical.
* Three outer iterations of all 24 volumes with the HIP-graph replay over 4
  streams (the first eager, then one capture replayed twice) against the
  same volumes advanced one at a time with eager launches: images and motion
  bitwise equal (the graph records the same kernels in the same per-volume
  order)."""
import os

import numpy as np
import pytest
import torch

from oracle import motion_oracle as O

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _motion_err(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    d = np.abs(a - np.asarray(b))
    return max(d[:, :3].max(), np.degrees(d[:, 3:]).max())


@pytest.fixture(scope="module")
def c4(me):
    O.set_workers(min(16, len(os.sched_getaffinity(0))))
    specs = me.simulation.batch_specs(24, seed=4)
    vols, host = [], []
    for sp in specs:
        d = me.simulation.make_volume(sp, device=DEV)
        kg, rkg = me.compute_transform_grids_voxel(sp["shape"], (1, 1, 1))
        vols.append(dict(ksp=d["ksp"], mps=d["mps"], shot_mask=d["A"], kgrid=kg, rkgrid=rkg))
    print(f"c4: {len(vols)} volumes, shapes {sorted({tuple(s['shape']) for s in specs})}", flush=True)
    return specs, vols


def test_c4_cheapest_volume_matches_oracle(me, c4):
    specs, vols = c4
    cost = [int(np.prod(sp["shape"])) * sp["coils"] * sp["nshots"] for sp in specs]
    i = int(np.argmin(cost))
    print(f"c4 oracle volume {i}: {specs[i]}", flush=True)
    bj = me.BatchedJointRecon(vols, nstreams=4)
    try:
        bj.step_eager()
        got_x, got_t = bj.recons[i].image().cpu().numpy(), bj.recons[i].t.cpu().numpy()
        got_damp = bj.recons[i].damp.cpu().numpy()
    finally:
        bj.close()
    del bj
    v = vols[i]
    shape = tuple(specs[i]["shape"])
    okg, orkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    ref = O.JointRecon(v["ksp"].cpu().numpy().astype(np.complex128), v["mps"].cpu().numpy().astype(np.complex128),
                       v["shot_mask"], okg, orkg, max_cg_iter=3, max_joint_iter=1, compute_objective=False)
    x_ref, t_ref = ref.run()
    assert float(np.linalg.norm(got_x - x_ref) / np.linalg.norm(x_ref)) < 1e-4
    assert _motion_err(got_t, t_ref) < 1e-3
    assert np.array_equal(got_damp, ref.damp)


def test_c4_graph_replay_equals_one_at_a_time(me, c4):
    specs, vols = c4
    n_iter = 3
    eager = me.BatchedJointRecon(vols, nstreams=4)
    for _ in range(n_iter):
        eager.step_eager()
    torch.cuda.synchronize()
    ref = [(r.image().cpu(), r.t.cpu()) for r in eager.recons]
    del eager
    torch.cuda.empty_cache()
    bj = me.BatchedJointRecon(vols, nstreams=4)
    try:
        imgs, ts = bj.run(n_iter)
        torch.cuda.synchronize()
        assert bj.steps_done == n_iter and bj._exec is not None
        moved = 0
        for (ri, rt), gi, gt in zip(ref, imgs, ts):
            assert torch.equal(gi.cpu(), ri)
            assert torch.equal(gt.cpu(), rt)
            moved += int(float(rt.abs().max()) > 0)
        assert moved == len(vols)
    finally:
        bj.close()
