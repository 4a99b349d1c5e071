"""The allocation-free forward encoding (me_encode_forward_static) against the
workspace one (me_encode_forward), called through the C ABI.

Both run the same kernels, so the gathered samples and the per-shot FP64
residual norms must agree bitwise.  Two calls in a row, the second on a
smaller problem (fewer shots, other extents), check that the shared static
partials array left by the first call does not leak into the second.  The
size limit is checked on the host before any device work."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _problem(me, shape, ncoils, tile, seed):
    sim = me.simulation
    gt = sim.phantom(shape, seed=seed).numpy()
    mps = sim.coil_maps(shape, ncoils, seed=seed + 1).numpy()
    A, _ = sim.disorder_sampling(shape, tile, 2)
    rng = np.random.default_rng(seed + 2)
    ksp = (rng.standard_normal((ncoils, *shape)) + 1j * rng.standard_normal((ncoils, *shape))).astype(np.complex64)
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    pb = me.encoding.SenseProblem(ksp, mps, A, kg, rkg)
    u = torch.stack([pb.image_to_plane(np.roll(gt, s, axis=0)) for s in range(pb.nshots)])
    return pb, u


def _both(me, pb, u):
    """(y, norms) of me_encode_forward with a workspace, then of the static variant."""
    L = me._lib
    pl = pb.plan
    dev = pb.device
    s = L.stream(dev)
    nps = L.size("me_encode_partials_size", pb.shape[2], pb.ncoils, pl.max_cnt, pb.shape[1])
    ws = torch.empty(nps * pb.nshots, dtype=torch.float64, device=dev)
    out = []
    for static in (False, True):
        y = torch.full_like(pb.ymeas, float("nan"))
        norms = torch.full((pb.nshots,), float("nan"), dtype=torch.float64, device=dev)
        common = (L.ptr(u), pb.vol, L.ptr(pb.maps), pb.ncoils, pb.nshots, L.int3(pb.shape), L.ptr(pl.sk0),
                  L.ptr(pl.sk1), L.ptr(pl.off), L.ptr(pl.cnt), pl.max_cnt, L.ptr(y), L.ptr(pb.ymeas))
        if static:
            L.call("me_encode_forward_static", *common, L.ptr(norms), 0, s)
        else:
            L.call("me_encode_forward", *common, L.ptr(ws), L.ptr(norms), 0, s)
        torch.cuda.synchronize()
        out.append((y.cpu().numpy(), norms.cpu().numpy()))
    return out


def test_static_forward_matches_workspace_forward(me):
    # first call: 32 x 24 x 16, 6 coils, 8 shots; second: 16 x 12 x 8, 3 coils, 4 shots
    for shape, ncoils, tile, seed in (((32, 24, 16), 6, (4, 4), 0), ((16, 12, 8), 3, (2, 4), 10)):
        pb, u = _problem(me, shape, ncoils, tile, seed)
        (y0, n0), (y1, n1) = _both(me, pb, u)
        print(f"static encode: shape {shape}, {ncoils} coils, {pb.nshots} shots", flush=True)
        assert np.isfinite(y0).all() and np.isfinite(n0).all()
        assert np.array_equal(y0, y1)
        assert np.array_equal(n0, n1)
        assert (n0 > 0).all()


def test_static_forward_without_norms_and_size_limit(me):
    pb, u = _problem(me, (16, 12, 8), 3, (2, 4), 20)
    L = me._lib
    pl = pb.plan
    y = torch.empty_like(pb.ymeas)
    ref = torch.empty_like(pb.ymeas)
    args = (L.ptr(u), pb.vol, L.ptr(pb.maps), pb.ncoils, pb.nshots, L.int3(pb.shape), L.ptr(pl.sk0), L.ptr(pl.sk1),
            L.ptr(pl.off), L.ptr(pl.cnt), pl.max_cnt)
    L.call("me_encode_forward_static", *args, L.ptr(y), None, None, 0, L.stream(pb.device))
    L.call("me_encode_forward", *args, L.ptr(ref), None, None, None, 0, L.stream(pb.device))
    torch.cuda.synchronize()
    assert torch.equal(y, ref)
    # a max_cnt whose partials exceed the static array is refused before any launch
    lib = L.load()
    huge = 1 << 28
    need = L.size("me_encode_partials_size", pb.shape[2], pb.ncoils, huge, pb.shape[1]) * pb.nshots
    assert need > (1 << 20)
    rc = lib.me_encode_forward_static(*args[:10], huge, L.ptr(y), L.ptr(pb.ymeas),
                                      L.ptr(torch.empty(pb.nshots, dtype=torch.float64, device=pb.device)), 0,
                                      L.stream(pb.device))
    assert rc != 0 and b"static scratch" in lib.me_last_error()
