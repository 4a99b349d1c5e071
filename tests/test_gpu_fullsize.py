"""Size-independent properties at the headline size (256 x 256 x 192,
32 coils, DISORDER tile 16 x 16, acceleration 2, 64 shots; 4 shots used),
where the FP64 oracle is too slow to run: unitarity and adjoint identities
of the transform, the encoding and the normal operator, and the round trip
T^H T = I; and the encoding and normal operator against a plain torch.fft
restatement of the reference's algorithm on the same GPU."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
SHAPE = (256, 256, 192)
NC = 32
NS = 4


def _rel(a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


@pytest.fixture(scope="module")
def full(me):
    dev = torch.device("cuda:0")
    sim = me.simulation
    A, _ = sim.disorder_sampling(SHAPE, (16, 16), 2, nshots=64)
    A = A[:NS]
    mps = sim.coil_maps(SHAPE, NC, seed=1, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    ksp = torch.zeros((NC, *SHAPE), dtype=torch.complex64, device=dev)
    kg, rkg = me.compute_transform_grids_voxel(SHAPE, (1, 1, 1))
    pb = me.encoding.SenseProblem(ksp, mps, A, kg, rkg, centered=True)
    del ksp, mps
    th = torch.as_tensor(sim.motion_uniform(NS, rot_deg=5.0, trans_vox=5.0, seed=8), device=dev)
    pb.theta.copy_(th)
    x = torch.randn(pb.vol, dtype=torch.complex64, device=dev, generator=g)
    y = torch.randn(pb.vol, dtype=torch.complex64, device=dev, generator=g)
    return dict(pb=pb, th=th, x=x, y=y, g=g, A=A)


def test_transform_unitary_and_round_trip(me, full):
    pb, th, x = full["pb"], full["th"], full["x"]
    U = torch.empty((NS, pb.vol), dtype=torch.complex64, device=x.device)
    pb._rigid(x, 0, U, th, 0)
    for s in range(NS):
        assert abs(float(torch.linalg.vector_norm(U[s]) / torch.linalg.vector_norm(x)) - 1) < 1e-5
    back = torch.empty_like(U)
    pb._rigid(U, pb.vol, back, th, 1)  # T^H
    for s in range(NS):
        assert _rel(back[s], x) < 1e-5


def test_transform_adjoint_identity(me, full):
    pb, th, x, y = full["pb"], full["th"], full["x"], full["y"]
    Tx = torch.empty((NS, pb.vol), dtype=torch.complex64, device=x.device)
    THy = torch.empty_like(Tx)
    pb._rigid(x, 0, Tx, th, 0)
    pb._rigid(y, 0, THy, th, 1)
    for s in range(NS):
        lhs = complex(torch.vdot(Tx[s], y))
        rhs = complex(torch.vdot(x, THy[s]))
        assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_encoding_adjoint_identity(me, full):
    pb, x = full["pb"], full["x"]
    U = torch.stack([x] * NS)
    Y = torch.empty_like(pb.ymeas)
    pb._encode(U, pb.vol, 0, NS, y=Y)
    z = torch.randn(Y.shape, dtype=torch.complex64, device=x.device, generator=full["g"])
    W = torch.empty((NS * pb.ncg, pb.vol), dtype=torch.complex64, device=x.device)
    pb._decode(z, 0, NS, W)
    W = W.reshape(NS, pb.ncg, pb.vol)[:, :pb.parts].sum(1)
    lhs = complex(torch.vdot(Y, z))
    rhs = sum(complex(torch.vdot(U[s], W[s])) for s in range(NS))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_normal_operator_hermitian_on_support(me, full):
    pb, x, y = full["pb"], full["x"], full["y"]
    xm, ym = pb.mask_(x.clone()), pb.mask_(y.clone())
    Ax, Ay = pb.zeros_image(), pb.zeros_image()
    pb.normal(xm, Ax)
    pb.normal(ym, Ay)
    lhs = complex(torch.vdot(Ax, ym))
    rhs = complex(torch.vdot(xm, Ay))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5
    assert float(torch.vdot(xm, Ax).real) > 0


def test_forward_and_normal_match_plain_torch_restatement(me, full):
    """The gathered encoding E_s T_s x and the normal operator against the
    reference's algorithm restated plainly with torch.fft on the same GPU
    (bench.py's comparison leg): 1-D FFT shears and a 3-D FFT translation,
    dense masks, 3-D FFTs of every coil.  complex64 both, relative L2 1e-4."""
    from bench import plain_transform

    pb, th, x = full["pb"], full["th"], full["x"]
    dev = x.device
    n2 = SHAPE[2]
    mps_u = torch.fft.ifftshift(me.simulation.coil_maps(SHAPE, NC, seed=1, device=dev), dim=(1, 2, 3))
    A_u = np.fft.ifftshift(full["A"], axes=(1, 2)).reshape(NS, SHAPE[0], SHAPE[1], 1)
    tkg, trkg = me.compute_transform_grids_voxel(SHAPE, (1, 1, 1), xp=torch)
    xi = pb.image_from_plane(x)
    theta = th.cpu().numpy()
    Y = pb.forward(x)
    pl = pb.plan
    ref_n = torch.zeros_like(xi)
    for s in range(NS):
        Tx = plain_transform(torch, xi, theta[s], tkg, trkg)
        k = torch.fft.fftn(mps_u * Tx, dim=(1, 2, 3), norm="ortho") * torch.as_tensor(A_u[s], device=dev)
        hyb = torch.fft.ifft(k, dim=3, norm="ortho")
        o, c = int(pl.off_h[s]), int(pl.cnt_h[s])
        k0 = torch.as_tensor(pl.k0_h[o:o + c], dtype=torch.long, device=dev)
        k1 = torch.as_tensor(pl.k1_h[o:o + c], dtype=torch.long, device=dev)
        got = Y[o * n2 * NC:(o + c) * n2 * NC].reshape(n2, NC, c).permute(1, 2, 0)
        assert _rel(got, hyb[:, k0, k1, :]) < 1e-4, s
        del hyb
        t = torch.fft.ifftn(k, dim=(1, 2, 3), norm="ortho")
        ref_n += plain_transform(torch, (mps_u.conj() * t).sum(0), theta[s], tkg, trkg, adjoint=True)
    ref_n *= (mps_u.abs().square().sum(0) > 0.1)
    out = pb.zeros_image()
    pb.normal(x, out)
    assert _rel(pb.image_from_plane(out), ref_n) < 1e-4


def test_espirit_routes_agree_at_full_size(me):
    """ESPIRiT at the full image size with 32 coils and the 5x5x5 kernel on a
    24^3 calibration region (m = 4000 rows, the reference's setting, where the
    FP64 oracle takes most of an hour): the subspace-iteration route and the
    full Jacobi route keep the same count and give the same maps."""
    from espirit_data import espirit_data
    calib, _, _ = espirit_data((48, 48, 48), nc=32, cal=(24, 24, 24), seed=1)
    calib = calib.astype(np.complex64)
    shape = (256, 256, 192)
    a = me.rx_espirit_3d(calib, shape, kernel_size=(5, 5, 5), method="subspace")
    ka = me.rx_espirit_3d.kept
    assert me.rx_espirit_3d.method == "subspace"
    b = me.rx_espirit_3d(calib, shape, kernel_size=(5, 5, 5), method="jacobi")
    assert me.rx_espirit_3d.kept == ka > 0
    ma, mb = (a != 0).any(-1), (b != 0).any(-1)
    assert float((ma != mb).float().mean()) < 1e-4  # mask ties at the threshold only
    both = ma & mb
    assert _rel(a[both], b[both]) < 1e-4
