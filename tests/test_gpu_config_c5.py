"""One outer iteration of the joint reconstruction at the headline size:
256 x 256 x 192 with 32 coils and 64 DISORDER shots (tile 16 x 16,
acceleration 2), k-space corrupted by 128 states of smooth cubic motion
plus 30 dB noise (simulation.make_config("c5")); 3 CG steps and one LM update
of all 64 x 6 parameters, then the objective.  The FP64 oracle cannot run
here in a test's time, so the library is compared with the plain torch.fft
restatement of the reference on the same GPU (bench.PlainJointRecon: dense
masks, coil-batched 3-D FFTs, a dense 6 x 32 x 256 x 256 x 192 Jacobian buffer,
27 dots per shot; held to the oracle in tests/test_plain_reference.py).
Relative L2 1e-4 on the image and the objective, 1e-3 (degrees, voxels) on
motion, LM damping identical."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_c5_outer_iteration_matches_plain_restatement(me):
    import bench

    dev = torch.device("cuda:0")
    d = me.simulation.make_config("c5", device=dev)
    shape = tuple(d["config"]["shape"])
    assert shape == (256, 256, 192) and d["config"]["coils"] == 32 and d["A"].shape[0] == 64
    print("c5: data ready", flush=True)
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    app = me.JointRecon(d["ksp"], d["mps"], d["A"], kg, rkg, max_cg_iter=3, max_joint_iter=1)
    x_got, t_got = app.run()
    x_got, t_got = x_got.cpu(), t_got.cpu().numpy()
    damp = app.damp.cpu().numpy()
    obj_got = float(app.objective_history[0])
    del app
    torch.cuda.empty_cache()
    print("c5: library iteration done", flush=True)
    tkg, trkg = me.compute_transform_grids_voxel(shape, (1, 1, 1), xp=torch)
    pj = bench.PlainJointRecon(torch, d["ksp"], d["mps"], d["A"], tkg, trkg, max_cg_iter=3)
    del d
    obj = pj.step()
    x_ref = pj.image().cpu()
    rel = float(torch.linalg.vector_norm(x_got - x_ref) / torch.linalg.vector_norm(x_ref))
    dt = np.abs(t_got - pj.t)
    merr = max(dt[:, :3].max(), np.degrees(dt[:, 3:]).max())
    print(f"c5: image rel {rel:.2e}, motion {merr:.2e}, objective rel {abs(obj_got - obj) / obj:.2e}, "
          f"LM accepted {sum(pj.accepted[0])}/64", flush=True)
    assert rel < 1e-4
    assert merr < 1e-3
    assert abs(obj_got - obj) / obj < 1e-4
    assert np.array_equal(damp, pj.damp)
