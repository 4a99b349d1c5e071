"""Encoding parity against the FP64 oracle on geometries that select the other
kernel variants (image_estimation.py:94-128, transform_estimation.py:42-45):

* column lengths N0 that are not powers of two (the adjoint's twiddle index
  by multiply-and-correct instead of a mask);
* rows longer than 256 (N1 = 384, 512: the row-group adjoint for lines of
  more than 16 threads);
* shots with more samples per readout plane than one forward workgroup holds
  (several sample chunks per plane) and more than the adjoint stages at once
  (512 per coil: several staged chunks);
* launch grids that are not a multiple of 8 workgroups (the workgroups are
  mapped onto the 8 XCDs in contiguous ranges, the first ones one longer).

Tolerance: relative L2 1e-4 on gathered k-space and images."""
import numpy as np
import pytest
import torch

from oracle import motion_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-4

# (shape, coils, tile, accel, shots): what each case exercises
CASES = [
    ((24, 12, 8), 4, (4, 4), 1, 2),    # N0 = 24, 130-134 samples > 128 per forward workgroup
    ((48, 32, 8), 5, (4, 4), 1, 2),    # N0 = 48, 768 samples: 3 forward chunks, 2 adjoint chunks
    ((96, 24, 8), 3, (8, 4), 2, 4),    # N0 = 96 = 3 * 32, 279-281 samples: 2 forward chunks
    ((16, 384, 8), 4, (4, 4), 2, 4),   # N1 = 384: row-group adjoint, power-of-two N0
    ((12, 512, 8), 3, (4, 4), 2, 4),   # N1 = 512, N0 = 12: row-group adjoint with multiply-and-correct
    ((32, 256, 8), 5, (4, 4), 1, 2),   # N1 = 256 (16-thread rows), 4096 samples: 8 forward chunks
    ((24, 12, 12), 3, (4, 4), 2, 3),   # 36-workgroup grids: not a multiple of the 8 XCDs
]


def _rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm(a - np.asarray(b)) / np.linalg.norm(b))


def _make_case(me, shape, nc, tile, accel, nshots):
    sim = me.simulation
    mps = sim.coil_maps(shape, nc, seed=21).numpy().astype(np.complex128)
    A, _ = sim.disorder_sampling(shape, tile, accel, nshots=nshots)
    t = sim.motion_uniform(A.shape[0], rot_deg=4.0, trans_vox=1.5, seed=22)
    rng = np.random.default_rng(23)
    ksp = rng.standard_normal((nc, *shape)) + 1j * rng.standard_normal((nc, *shape))
    A_u = np.fft.ifftshift(A, axes=(1, 2))
    mps_u = np.fft.ifftshift(mps, axes=(-3, -2, -1))
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    okg, orkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    pb = me.encoding.SenseProblem(ksp, mps_u, A_u, kg, rkg)
    pb.theta.copy_(torch.as_tensor(t))
    M = np.sum(np.abs(mps_u) ** 2, axis=0) > 0.1
    return dict(shape=shape, nc=nc, pb=pb, mps_u=mps_u, A_u=A_u, ksp=ksp, t=t, okg=okg, orkg=orkg, M=M, rng=rng)


@pytest.fixture(scope="module", params=CASES, ids=lambda c: "x".join(map(str, c[0])))
def case(request, me):
    return _make_case(me, *request.param)


def test_case_selects_the_intended_paths(case):
    """The sample counts that make each case cover several chunks."""
    cnt = int(case["pb"].plan.cnt_h.max())
    need = {(24, 12, 8): 129, (48, 32, 8): 513, (96, 24, 8): 257, (16, 384, 8): 513, (12, 512, 8): 513,
            (32, 256, 8): 4096, (24, 12, 12): 48}
    assert cnt >= need[case["shape"]], cnt


def test_forward_encoding(case):
    d = case
    pb, shape, nc = d["pb"], d["shape"], d["nc"]
    x = d["rng"].standard_normal(shape) + 1j * d["rng"].standard_normal(shape)
    Y = pb.forward(pb.image_to_plane(x)).cpu().numpy()
    pl = pb.plan
    n2 = shape[2]
    for s in range(pb.nshots):
        dense = O.encode_forward(x, d["mps_u"], d["A_u"][s], d["t"][s], d["okg"], d["orkg"])
        hyb = np.fft.ifft(dense, axis=-1, norm="ortho")
        o, c = pl.off_h[s], pl.cnt_h[s]
        ref = hyb[:, pl.k0_h[o:o + c], pl.k1_h[o:o + c], :]
        got = Y[o * n2 * nc:(o + c) * n2 * nc].reshape(n2, nc, c)
        assert _rel(np.transpose(got, (1, 2, 0)), ref) < TOL, s


def test_normal_operator_and_rhs(case):
    d = case
    pb, shape = d["pb"], d["shape"]
    x = d["rng"].standard_normal(shape) + 1j * d["rng"].standard_normal(shape)
    out = pb.zeros_image()
    pb.normal(pb.image_to_plane(x), out)
    ref = O.normal_op(x, d["mps_u"], d["A_u"], d["t"], d["okg"], d["orkg"], d["M"])
    assert _rel(pb.image_from_plane(out), ref) < TOL
    refb = O.cg_rhs(d["ksp"], d["mps_u"], d["A_u"], d["t"], d["okg"], d["orkg"], d["M"])
    assert _rel(pb.image_from_plane(pb.rhs(cache=False)), refb) < TOL
    assert _rel(pb.image_from_plane(pb.rhs(cache=True)), refb) < TOL


def test_forward_variants_bitwise_equal_and_at_parity(me, case):
    """Every forward-encoding variant (16 coils x 2 samples per thread, 8 x 2,
    8 x 4, 4 x 4; me_encode_tuning_set) gives bitwise the same samples, residual
    norms equal to FP64 rounding, and the oracle's encoding."""
    d = case
    pb, shape, nc = d["pb"], d["shape"], d["nc"]
    L = me._lib
    x = pb.image_to_plane(d["rng"].standard_normal(shape) + 1j * d["rng"].standard_normal(shape))
    key = (L.int3(shape), nc, pb.plan.max_cnt)
    outs = []
    try:
        for v in (0, 1, 2, 3):
            L.call("me_encode_tuning_set", *key, v)
            assert L.size("me_encode_tuning_get", *key) == v
            Y = pb.forward(x).clone()
            norms = pb.objective_per_shot(x)
            outs.append((Y, norms))
    finally:
        L.load().me_encode_tuning_clear()
    assert L.size("me_encode_tuning_get", *key) == 0
    for Y, norms in outs[1:]:
        assert torch.equal(Y, outs[0][0])
        assert float((norms - outs[0][1]).abs().max() / outs[0][1].abs().max()) < 1e-12
    pl = pb.plan
    n2 = shape[2]
    xh = pb.image_from_plane(x).cpu().numpy()
    Y = outs[2][0].cpu().numpy()
    dense = O.encode_forward(xh, d["mps_u"], d["A_u"][0], d["t"][0], d["okg"], d["orkg"])
    hyb = np.fft.ifft(dense, axis=-1, norm="ortho")
    c = pl.cnt_h[0]
    ref = hyb[:, pl.k0_h[:c], pl.k1_h[:c], :]
    assert _rel(np.transpose(Y[:c * n2 * nc].reshape(n2, nc, c), (1, 2, 0)), ref) < TOL


def test_forward_encoder_tuner_records_its_pick(me, case):
    import ctypes as C

    d = case
    pb, shape, nc = d["pb"], d["shape"], d["nc"]
    L = me._lib
    try:
        picks = me.batched.tune_encoders(pb, reps=2)
        (variant, ms), = picks.values()
        assert variant in (0, 1, 2, 3) and min(ms) > 0
        assert ms[variant] == min(ms)
        assert L.size("me_encode_tuning_get", L.int3(shape), nc, pb.plan.max_cnt) == variant
    finally:
        L.load().me_encode_tuning_clear()


def test_whole_and_partial_chunk_forms_bitwise_equal(me):
    """The forward encoding accumulates its sample slots without a per-lane
    test when max_cnt fills whole chunks and with one otherwise; the same
    shots encoded under both forms (max_cnt as is, 4096, a multiple of every
    variant's chunk, then one more, which adds an empty chunk) give bitwise
    the same samples and residual norms."""
    d = _make_case(me, (32, 256, 8), 5, (4, 4), 1, 2)
    pb, shape, nc = d["pb"], d["shape"], d["nc"]
    L = me._lib
    x = pb.image_to_plane(d["rng"].standard_normal(shape) + 1j * d["rng"].standard_normal(shape))
    U = torch.empty((pb.nshots, pb.vol), dtype=torch.complex64, device=pb.device)
    pb._rigid(x, 0, U, pb.theta, 0)
    pl = pb.plan
    maxc = int(pl.cnt_h.max())
    outs = []
    for m in (maxc, maxc + 1):
        Y = torch.zeros_like(pb.ymeas)
        nps = int(L.size("me_encode_partials_size", shape[2], nc, m, shape[1]))
        part = torch.empty(max(1, nps * pb.nshots), dtype=torch.float64, device=pb.device)
        norms = torch.empty(pb.nshots, dtype=torch.float64, device=pb.device)
        L.call("me_encode_forward", L.ptr(U), pb.vol, L.ptr(pb.maps), nc, pb.nshots, L.int3(shape), L.ptr(pl.sk0),
               L.ptr(pl.sk1), L.ptr(pl.off), L.ptr(pl.cnt), m, L.ptr(Y), L.ptr(pb.ymeas), L.ptr(part),
               L.ptr(norms), 1, L.stream(pb.device))
        outs.append((Y, norms))
    assert torch.equal(outs[0][0], outs[1][0])
    assert float((outs[0][1] - outs[1][1]).abs().max() / outs[0][1].abs().max()) < 1e-12
