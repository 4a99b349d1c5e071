#!/usr/bin/env python3
"""Benchmark: joint-reconstruction outer iterations per second on one MI355X.

Workload (default): 256 x 256 x 192 complex64 phantom, 32 coils, 64 DISORDER
shots (tile 16 x 16, acceleration 2), k-space corrupted by 128 motion states
of smooth cubic trajectories (+-5 deg, +-5 voxels) plus 30 dB noise, all
synthetic and seeded.  One step = one outer iteration of JointRecon: right-hand
side, initial residual + 3 CG iterations, one LM update of all 64 x 6
parameters with mean re-registration, x *= M, and one objective value (that
of the previous iteration's final state, taken from the CG's first encoding
of that same state).

Prints ONE JSON line with the throughput, a roofline entry for the dominant
kernel (timed live with HIP events on the stream it runs on) and a CPU
baseline (the complex64 NumPy/SciPy restatement timed on a bounded sample).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c5", help="c1, c2, c3 or c5 (default: the headline workload)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-objective", action="store_true", help="skip the per-iteration objective")
    ap.add_argument("--no-tune", action="store_true", help="skip the line-pass autotuner at setup")
    ap.add_argument("--grad", action="store_true",
                    help="differentiation workload instead of the headline: data-consistency loss forward+backward "
                         "at c3 (192x192x128, 24 coils, 48 shots), recompute on/off, and a 64-direction central "
                         "difference check at step 1e-3 (prints its own JSON line)")
    ap.add_argument("--batch", action="store_true",
                    help="batching workload instead of the headline: 24 seeded volumes of uneven size "
                         "(64x64x48 / 96x96x64 / 128x128x96), coils 8-20, shots 16-64; one-at-a-time eager outer "
                         "iterations against graph replays over concurrent streams, before and after the "
                         "line-pass autotuner (prints its own JSON line)")
    ap.add_argument("--calib", action="store_true",
                    help="calibration workload instead of the headline: noise whitening, ESPIRiT (24^3 "
                         "calibration region, kernel 5^3) and coil compression (perc 0.95) at 256x256x192 with 32 "
                         "coils; per-stage GPU times and a CPU baseline on bounded samples (prints its own JSON line)")
    ap.add_argument("--plain", action="store_true",
                    help="comparison workload instead of the headline: the normal operator (sum over 64 shots of "
                         "T^H E^H E T) and the forward E T at 256x256x192 with 32 coils, through the library and "
                         "through a plain torch.fft restatement of the reference's algorithm on the same GPU (dense "
                         "masks, 3-D FFTs of every coil, 1-D FFT shears); prints its own JSON line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (three times), the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE to bytes")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_encode_fwd_256x256x192_c32_s64.md")


def read_traffic():
    """HBM bytes per k_encode_fwd launch measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE (scaled by a plain copy), recorded in the profile summary."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def copy_calibration(torch):
    n = 1 << 30
    a = torch.ones(n // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 3}), flush=True)


class KernelTimer:
    """Per-entry-point HIP event timing (events recorded on torch's current
    stream, the stream every library call is enqueued on)."""

    def __init__(self, torch):
        self.torch = torch
        self.events = {}

    def wrap(self, name, fn):
        torch = self.torch

        def inner(*a, **k):
            s = torch.cuda.Event(enable_timing=True)
            e = torch.cuda.Event(enable_timing=True)
            s.record()
            out = fn(*a, **k)
            e.record()
            self.events.setdefault(name, []).append((s, e, a, k))
            return out

        return inner

    def durations(self, name):
        return [s.elapsed_time(e) * 1e-3 for s, e, _, _ in self.events.get(name, [])]


def encode_bytes(pb, s0, s1, meas):
    """Algorithmic bytes of one me_encode_forward launch over shots s0..s1:
    per shot, the image (V) and the coil maps (C*V) read once, the gathered
    samples written (8*C*K_s) and, for a residual, the measured samples read."""
    V = pb.vol * 8
    cnt = pb.plan.cnt_h[s0:s1]
    K = cnt * pb.shape[2]
    per = V + pb.ncoils * V + 8 * pb.ncoils * K * (2 if meas else 1)
    return float(per.sum())


def decode_bytes(pb, s0, s1):
    """Algorithmic bytes of one me_encode_adjoint launch over shots s0..s1: per
    shot, the gathered samples (8*C*K_s) and the coil maps (C*V) read once and
    one image (V) written."""
    V = pb.vol * 8
    K = pb.plan.cnt_h[s0:s1] * pb.shape[2]
    return float((V + pb.ncoils * V + 8 * pb.ncoils * K).sum())


def transform_bytes(pb, B):
    """Algorithmic bytes of one forward rigid transform of a single image
    broadcast to B shots (me_rigid_transform mode 0, shared input, 10 line
    passes): the first pass reads the image once and writes B volumes, the
    other nine read and write B volumes each."""
    return float(pb.vol * 8 * (19 * B + 1))


def cpu_baseline(shape, ncoils, nshots, n_cg):
    """Times the oracle's complex64 restatement (dense masks, 3-D FFTs) on one
    sample of each unit at the full shape and composes one outer iteration."""
    import numpy as np

    from oracle import motion_oracle as O

    cores = len(os.sched_getaffinity(0))
    cores = min(cores, 16)
    O.set_workers(cores)
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    mps = np.ones((ncoils, *shape), dtype=np.complex64) / np.sqrt(ncoils)
    mask = np.zeros((shape[0], shape[1], 1), dtype=bool)
    mask[::16, ::16] = True
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    th = np.array([0.5, -0.3, 0.2, 0.02, -0.01, 0.015])
    t = {}
    t0 = time.perf_counter()
    Tx = O.rigid_apply(x, th, kg, rkg, prec="c64")
    t["T"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.rigid_deriv(x, th, kg, rkg, 4, prec="c64")
    t["dT"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    k = O._fftn((mps * Tx).astype(np.complex64), (-3, -2, -1), O.WORKERS) * mask
    t["Ecore"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    tmp = O._ifftn((mask * k).astype(np.complex64), (-3, -2, -1), O.WORKERS)
    np.sum(np.conj(mps) * tmp, axis=0)
    t["EHcore"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    np.vdot(k, tmp)
    t["dot"] = time.perf_counter() - t0
    total_sample = sum(t.values())
    E = t["T"] + t["Ecore"]
    EH = t["EHcore"] + t["T"]
    outer = nshots * (EH                              # right-hand side
                      + (n_cg + 1) * (E + EH)          # initial residual + CG steps
                      + 2 * E + 6 * (t["dT"] + t["Ecore"]) + 27 * t["dot"]  # LM (residual, 6 columns, candidate)
                      + E)                             # objective
    outer += t["T"]  # mean re-registration
    return dict(value=1.0 / outer, unit="outer_iters/s", cores=cores, kind="port",
                sample=(f"oracle complex64 restatement at {shape[0]}x{shape[1]}x{shape[2]}, {ncoils} coils: one each of "
                        f"T ({t['T']:.2f}s), dT ({t['dT']:.2f}s), coil-batched forward ({t['Ecore']:.2f}s) and adjoint "
                        f"({t['EHcore']:.2f}s) encoding, dense vdot ({t['dot']:.2f}s) = {total_sample:.1f}s, composed "
                        f"into one outer iteration ({nshots} shots, {n_cg} CG, 27 LM dots/shot) = {outer:.0f}s"))


def grad_bench(args, torch, me, dev):
    """Forward+backward of the data-consistency loss at c3 with recomputed or
    stored sheared intermediates, plus a central-difference check."""
    import numpy as np

    data = me.simulation.make_config("c3", device=dev)
    cfg = data["config"]
    shape, ncoils, nshots = cfg["shape"], cfg["coils"], cfg["nshots"]
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    pb = me.encoding.SenseProblem(data["ksp"], data["mps"], data["A"], kg, rkg, centered=True)
    x0 = torch.fft.ifftshift(data["gt"].to(dev), dim=(0, 1, 2)).to(torch.complex64)
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)
    # evaluation point: the true image with the motion not yet estimated
    # (small random parameters), where the loss is far from its minimum
    t0 = 0.01 * torch.randn((nshots, 6), dtype=torch.float32, device=dev, generator=gen)
    del data

    def fwd_bwd(recompute):
        x = x0.clone().requires_grad_(True)
        th = t0.clone().requires_grad_(True)
        me.data_consistency_loss(x, th, pb, recompute=recompute).backward()
        return x.grad, th.grad

    res = {}
    for name, fn in (("forward", lambda: me.data_consistency_loss(x0, t0, pb)),
                     ("fwd_bwd_recompute", lambda: fwd_bwd(True)),
                     ("fwd_bwd_stored", lambda: fwd_bwd(False))):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t_start) * 1e3 / args.steps
        log(f"{name}: {res[name]:.2f} ms")
    gx, gt = fwd_bwd(True)
    gnx, gnt = float(torch.linalg.vector_norm(gx)), float(torch.linalg.vector_norm(gt))
    rel, scaled = [], []
    h = 1e-3
    with torch.no_grad():
        for _ in range(64):
            dx = torch.randn(shape, dtype=torch.complex64, device=dev, generator=gen)
            dt = torch.randn((nshots, 6), dtype=torch.float32, device=dev, generator=gen)
            fd = (float(me.data_consistency_loss(x0 + h * dx, t0 + h * dt, pb))
                  - float(me.data_consistency_loss(x0 - h * dx, t0 - h * dt, pb))) / (2 * h)
            an = float(torch.vdot(gx.flatten(), dx.flatten()).real) + float((gt * dt).sum())
            rel.append(abs(fd - an) / abs(an))
            scaled.append(abs(fd - an) / (gnx * float(torch.linalg.vector_norm(dx))
                                          + gnt * float(torch.linalg.vector_norm(dt))))
    out = {
        "metric": "data-consistency loss forward+backward per second at 192x192x128, 24 coils, 48 shots",
        "value": round(1e3 / res["fwd_bwd_recompute"], 3), "unit": "fwd+bwd/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "higher_is_better": True, "dtype": "c64", "data": "synthetic",
        "config": {"workload": f"c3: {shape[0]}x{shape[1]}x{shape[2]} complex64, {ncoils} coils, {nshots} "
                               f"DISORDER shots, gradients w.r.t. image and 48x6 motion"},
        "ms": {k: round(v, 3) for k, v in res.items()},
        "fd_check": {"directions": 64, "step": h, "median_rel": float(np.median(rel)), "max_rel": float(max(rel)),
                     "max_rel_to_gradient_scale": float(max(scaled))},
    }
    print(json.dumps(out), flush=True)


def plain_transform(torch, x, th, kg, rkg, adjoint=False):
    """The reference's FFT-shear rigid transform (transform.py:263-377) written
    plainly with torch.fft: rotations about x, y, z of three 1-D Fourier shears
    each (tan, sin, tan), then a 3-D Fourier translation; the adjoint applies
    the conjugated translation first and the conjugated shears in reverse."""
    q1, q2, q3, t1, t2, t3 = [float(v) for v in th]
    steps = [(t1, rkg["tan"]["x"], rkg["sin"]["x"], 1, 2), (t2, rkg["tan"]["y"], rkg["sin"]["y"], 2, 0),
             (t3, rkg["tan"]["z"], rkg["sin"]["z"], 0, 1)]
    sgn = -1.0 if adjoint else 1.0

    def shear(a, ax, ph):
        return torch.fft.ifft(torch.fft.fft(a, dim=ax) * ph, dim=ax)

    def translate(a):
        U = torch.exp(-1j * sgn * (q1 * kg["x"] + q2 * kg["y"] + q3 * kg["z"])).to(torch.complex64)
        return torch.fft.ifftn(torch.fft.fftn(a) * U)

    if adjoint:
        x = translate(x)
    for th_s, rkt, rks, at, as_ in (steps[::-1] if adjoint else steps):
        vt = torch.exp(1j * sgn * math.tan(th_s / 2) * rkt).to(torch.complex64)
        vs = torch.exp(-1j * sgn * math.sin(th_s) * rks).to(torch.complex64)
        x = shear(shear(shear(x, at, vt), as_, vs), at, vt)
    return x if adjoint else translate(x)


def plain_deriv(torch, x, th, kg, rkg, p_idx):
    """RigidTransformDerivativeCuda.apply (transform.py:379-504) written
    plainly with torch.fft: rotations about z, y, x (the reverse of apply),
    the three-term product rule over the shears of the differentiated
    rotation (transform.py:433-474), then the translation, or for p < 3 its
    derivative -i k_p U (transform.py:422-431)."""
    q1, q2, q3, t1, t2, t3 = [float(v) for v in th]
    c64 = torch.complex64

    def fft(a, ax):
        return torch.fft.fft(a, dim=ax, norm="ortho")

    def ifft(a, ax):
        return torch.fft.ifft(a, dim=ax, norm="ortho")

    def phases(theta, rkt, rks):
        return (torch.exp((1j * math.tan(theta / 2) * rkt).to(c64)),
                torch.exp((-1j * math.sin(theta) * rks).to(c64)))

    def rot(a, theta, rkt, rks, at, as_):
        vt, vs = phases(theta, rkt, rks)
        a = ifft(fft(a, at) * vt, at)
        a = ifft(fft(a, as_) * vs, as_)
        return ifft(fft(a, at) * vt, at)

    def rot_d(a, theta, rkt, rks, at, as_):
        vt, vs = phases(theta, rkt, rks)
        vt_d = (1j * ((1 + math.tan(theta / 2) ** 2) / 2) * rkt * vt).to(c64)
        vs_d = (-1j * math.cos(theta) * rks * vs).to(c64)
        a = fft(a, at)
        o = fft(ifft(fft(ifft(a * vt_d, at), as_) * vs, as_), at) * vt
        s2 = fft(ifft(fft(ifft(a * vt, at), as_) * vs_d, as_), at) * vt
        s3 = fft(ifft(fft(ifft(a * vt, at), as_) * vs, as_), at) * vt_d
        return ifft(o + s2 + s3, at)

    def trans(a, p=None):
        U = torch.exp(-1j * (q1 * kg["x"] + q2 * kg["y"] + q3 * kg["z"]))
        if p is not None:
            U = -1j * kg["xyz"[p]] * U
        return torch.fft.ifftn(torch.fft.fftn(a, norm="ortho") * U.to(c64), norm="ortho")

    Z = (t3, rkg["tan"]["z"], rkg["sin"]["z"], 0, 1)
    Y = (t2, rkg["tan"]["y"], rkg["sin"]["y"], 2, 0)
    X = (t1, rkg["tan"]["x"], rkg["sin"]["x"], 1, 2)
    if p_idx < 3:
        return trans(rot(rot(rot(x, *Z), *Y), *X), p_idx)
    if p_idx == 3:
        return trans(rot_d(rot(rot(x, *Z), *Y), *X))
    if p_idx == 4:
        return trans(rot(rot_d(rot(x, *Z), *Y), *X))
    return trans(rot(rot(rot_d(x, *Z), *Y), *X))


class PlainJointRecon:
    """JointRecon (joint_recon.py:38-115) over estimate_image_cg
    (image_estimation.py:74-165) and estimate_transform
    (transform_estimation.py:36-152), restated plainly with torch.fft on the
    GPU as the reference's CuPy path runs: dense boolean masks, coil-batched
    3-D FFTs of every coil, a dense (6, C, N0, N1, N2) Jacobian buffer, 27 dot
    products per shot (FP32, as cupy.vdot on complex64), the 6x6 solve in FP64
    on the host, a host synchronisation per dot as the reference's .item() /
    scalar reads.  Inputs centred, as the reference's JointRecon takes them."""

    def __init__(self, torch, ksp, mps, shot_mask, kgrid, rkgrid, t0=None, max_cg_iter=3, tol=1e-6):
        import numpy as np

        self.torch = torch
        dims = (-3, -2, -1)
        dev = ksp.device
        self.kg, self.rkg = kgrid, rkgrid
        self.max_cg_iter, self.tol = max_cg_iter, tol
        self.ksp = torch.fft.ifftshift(ksp.to(torch.complex64), dim=dims)
        self.mps = torch.fft.ifftshift(mps.to(torch.complex64), dim=dims)
        self.mask = torch.fft.ifftshift(torch.as_tensor(np.asarray(shot_mask), device=dev), dim=dims)
        self.max_norm = float(torch.fft.ifftn(self.ksp, dim=dims, norm="ortho").abs().max())
        self.ksp /= self.max_norm
        self.S = self.mask.shape[0]
        self.x = torch.zeros(self.ksp.shape[1:], dtype=torch.complex64, device=dev)
        self.t = np.zeros((self.S, 6)) if t0 is None else np.array(t0, dtype=np.float64)
        self.damp = np.ones(self.S)
        self.flags = np.zeros(self.S, dtype=bool)
        self.partial = torch.empty((6, *self.ksp.shape), dtype=torch.complex64, device=dev)
        ssq = (self.mps.abs() ** 2).sum(0)
        self.P = 1 / (ssq + 1e-3)
        self.M = ssq > 0.1
        self.accepted = []

    def _T(self, v, th, adjoint=False):
        return plain_transform(self.torch, v, th, self.kg, self.rkg, adjoint=adjoint)

    def _fftn(self, a):
        return self.torch.fft.fftn(a, dim=(-3, -2, -1), norm="ortho")

    def _ifftn(self, a):
        return self.torch.fft.ifftn(a, dim=(-3, -2, -1), norm="ortho")

    def _vdot(self, a, b):
        return float(self.torch.vdot(a.reshape(-1), b.reshape(-1)).real)

    def normal(self, v):
        """A_func, image_estimation.py:106-129."""
        out = self.torch.zeros_like(v)
        for i in range(self.S):
            Ex = self._ifftn(self._fftn(self.mps * self._T(v, self.t[i])) * self.mask[i])
            out += self._T((self.mps.conj() * Ex).sum(0), self.t[i], adjoint=True)
        return out * self.M

    def cg(self):
        """estimate_image_cg, image_estimation.py:92-165."""
        b = self.torch.zeros_like(self.x)
        for i in range(self.S):
            tmp = self._ifftn(self.mask[i] * self.ksp)
            b += self._T((self.mps.conj() * tmp).sum(0), self.t[i], adjoint=True)
            b *= self.M
        x = self.x.clone()
        r = b - self.normal(x)
        z = self.P * r
        p = z.clone()
        rsold = self._vdot(r, z)
        for _ in range(self.max_cg_iter):
            Ap = self.normal(p)
            alpha = rsold / self._vdot(p, Ap)
            x = x + alpha * p
            r = r - alpha * Ap
            z = self.P * r
            rsnew = self._vdot(r, z)
            if math.sqrt(abs(rsnew)) < self.tol:
                break
            p = z + (rsnew / rsold) * p
            rsold = rsnew
        self.x = x

    def _residual(self, th):
        return (self._fftn(self.mps * self._T(self.x, th)) - self.ksp) * self.mask_s

    def lm(self):
        """estimate_transform, transform_estimation.py:39-152 (use_ga=False)."""
        import numpy as np

        acc = []
        for s in range(self.S):
            self.mask_s = self.mask[s]
            res = self._residual(self.t[s])
            e0 = self._vdot(res, res)
            g = np.zeros(6)
            for p in range(6):
                self.partial[p] = self._fftn(self.mps * plain_deriv(self.torch, self.x, self.t[s], self.kg, self.rkg,
                                                                    p)) * self.mask_s
                g[p] = self._vdot(res, self.partial[p])
            H = np.zeros((6, 6))
            for i in range(6):
                for j in range(i, 6):
                    v = self._vdot(self.partial[i], self.partial[j])
                    H[i, j] = v + self.damp[s] if i == j else v
                    H[j, i] = H[i, j]
            nxt = self.t[s] - np.linalg.solve(H, g)
            res = self._residual(nxt)
            e1 = self._vdot(res, res)
            if e1 < e0:
                if np.all(np.abs(self.t[s] - nxt) < self.tol):
                    self.flags[s] = True
                self.t[s] = nxt
                self.damp[s] = max(self.damp[s] / 5, 1e-4)
            else:
                self.damp[s] = min(self.damp[s] * 1.5, 1e16)
            acc.append(bool(e1 < e0))
        self.accepted.append(acc)
        mean = self.t.mean(0)
        self.x = self._T(self.x, mean)
        self.t -= mean

    def objective(self):
        """joint_recon.py:108-115."""
        total = 0.0
        for s in range(self.S):
            self.mask_s = self.mask[s]
            r = self._residual(self.t[s])
            total += self._vdot(r, r)
        return total

    def step(self):
        """One outer iteration, joint_recon.py:93-102."""
        self.cg()
        self.lm()
        self.x = self.x * self.M
        return self.objective()

    def image(self):
        return self.torch.fft.fftshift(self.x * self.max_norm)


def plain_bench(args, torch, me, dev):
    """Library against a plain torch.fft restatement of the reference's
    per-shot loop (image_estimation.py:106-129: transform, coil maps, 3-D FFT
    of every coil, dense mask, inverse 3-D FFT, conj(maps) coil sum, adjoint
    transform) on the same GPU, same inputs, complex64 both."""
    import numpy as np

    cfg = me.simulation.CONFIGS["c5"]
    shape, ncoils, nshots = cfg["shape"], cfg["coils"], cfg["nshots"]
    mps = me.simulation.coil_maps(shape, ncoils, seed=1, device=dev)
    A, _ = me.simulation.disorder_sampling(shape, cfg["tile"], cfg["accel"], nshots=nshots)
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    pb = me.encoding.SenseProblem(torch.zeros_like(mps), mps, A, kg, rkg, centered=True)
    rng = np.random.default_rng(11)
    theta = np.concatenate([rng.uniform(-2, 2, (nshots, 3)), np.radians(rng.uniform(-3, 3, (nshots, 3)))], 1)
    pb.theta.copy_(torch.as_tensor(theta))
    gen = torch.Generator(device=dev)
    gen.manual_seed(12)
    x = torch.randn(shape, dtype=torch.complex64, device=dev, generator=gen)
    # the plain path's operands in the reference's uncentred layout
    mps_u = torch.fft.ifftshift(mps, dim=(1, 2, 3)).contiguous()
    del mps
    masks = torch.as_tensor(np.fft.ifftshift(A, axes=(1, 2)), device=dev).reshape(nshots, shape[0], shape[1], 1)
    M = (mps_u.abs().square().sum(0) > 0.1).to(torch.complex64)
    tkg, trkg = me.compute_transform_grids_voxel(shape, (1, 1, 1), xp=torch)

    def plain_forward():
        k = None
        for s in range(nshots):
            Tx = plain_transform(torch, x, theta[s], tkg, trkg)
            k = torch.fft.fftn(mps_u * Tx, dim=(1, 2, 3), norm="ortho") * masks[s]
        return k

    def plain_normal():
        out = torch.zeros_like(x)
        for s in range(nshots):
            Tx = plain_transform(torch, x, theta[s], tkg, trkg)
            k = torch.fft.fftn(mps_u * Tx, dim=(1, 2, 3), norm="ortho") * masks[s]
            t = torch.fft.ifftn(k, dim=(1, 2, 3), norm="ortho")
            out += plain_transform(torch, (mps_u.conj() * t).sum(0), theta[s], tkg, trkg, adjoint=True)
        return out * M

    xp = pb.image_to_plane(x)
    outp = pb.zeros_image()

    def ours_normal():
        outp.zero_()
        return pb.normal(xp, outp)

    res = {}
    for name, fn in (("ours_forward", lambda: pb.forward(xp)), ("ours_normal", ours_normal),
                     ("plain_forward", plain_forward), ("plain_normal", plain_normal)):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t_start) * 1e3 / args.steps
        log(f"{name}: {res[name]:.2f} ms")
    ref = plain_normal()
    got = pb.image_from_plane(ours_normal())
    rel = float(torch.linalg.vector_norm(got - ref) / torch.linalg.vector_norm(ref))
    out = {
        "metric": "normal operator applications per second at 256x256x192, 32 coils, 64 shots",
        "value": round(1e3 / res["ours_normal"], 3), "unit": "applications/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "higher_is_better": True, "dtype": "c64", "data": "synthetic",
        "config": {"workload": f"{shape[0]}x{shape[1]}x{shape[2]} complex64, {ncoils} coils, {nshots} DISORDER "
                               f"shots, random image and motion (+-2 voxels, +-3 degrees)"},
        "ms": {k: round(v, 2) for k, v in res.items()},
        "speedup_over_plain": {"normal": round(res["plain_normal"] / res["ours_normal"], 2),
                               "forward": round(res["plain_forward"] / res["ours_forward"], 2)},
        "normal_rel_l2_vs_plain": rel,
    }
    print(json.dumps(out), flush=True)


def batch_bench(args, torch, me, dev):
    """Batched vs one-at-a-time joint reconstruction of uneven volumes."""
    specs = me.simulation.batch_specs(24, seed=4)
    vols = []
    t_setup = time.perf_counter()
    for i, sp in enumerate(specs):
        d = me.simulation.make_volume(sp, device=dev)
        kg, rkg = me.compute_transform_grids_voxel(sp["shape"], (1, 1, 1))
        vols.append(dict(ksp=d["ksp"], mps=d["mps"], shot_mask=d["A"], kgrid=kg, rkgrid=rkg))
        del d
    log(f"generated {len(vols)} volumes in {time.perf_counter() - t_setup:.1f}s")
    bj = me.BatchedJointRecon(vols, nstreams=4)
    del vols
    me._lib.load().me_line_tuning_clear()
    bj.step_eager()  # allocates every workspace
    torch.cuda.synchronize()
    log("warm-up outer iteration done")
    nvol = len(bj.recons)
    res = {}

    def timed(name, fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[name] = dict(ms_per_iteration=round(dt * 1e3 / args.steps, 2),
                         volume_iters_per_s=round(nvol * args.steps / dt, 2))
        log(f"{name}: {res[name]}")

    timed("one_at_a_time", bj.step_eager)
    timed("graph_default_lines", bj.step)
    bj.close()
    tuning = bj.tune(reps=5)
    log(f"tuning: {tuning}")
    timed("graph_tuned", bj.step)
    bj.close()
    me._lib.load().me_line_tuning_clear()
    voxels = sum(int(r.problem.vol) for r in bj.recons)
    out = {
        "metric": "batched joint-recon volume outer-iterations/s, 24 uneven volumes",
        "value": res["graph_tuned"]["volume_iters_per_s"], "unit": "volume_iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": "24 volumes, shapes from {64x64x48, 96x96x64, 128x128x96}, coils 8-20, DISORDER "
                               "shots 16-64, random-walk motion, 3 CG + 1 LM per outer iteration",
                   "volumes": [dict(shape=list(sp["shape"]), coils=sp["coils"], shots=sp["nshots"],
                                    tile=list(sp["tile"]), accel=sp["accel"]) for sp in specs],
                   "total_voxels": voxels, "streams": len(bj.streams)},
        "legs": res,
        "speedup_vs_one_at_a_time": round(res["graph_tuned"]["volume_iters_per_s"]
                                          / res["one_at_a_time"]["volume_iters_per_s"], 3),
        "tuning": {"x".join(map(str, k)): dict(lines=list(v[0]), ms=list(v[1])) for k, v in tuning.items()},
    }
    print(json.dumps(out), flush=True)


def calib_bench(args, torch, me, dev):
    """Calibration chain of recon.py:130-150 on the GPU, per stage, with the
    oracle timed on bounded samples of the same inputs."""
    import numpy as np

    shape, C = (256, 256, 192), 32
    rng = np.random.default_rng(9)
    mps = me.simulation.coil_maps(shape, C, seed=1, device=dev)
    gt = me.simulation.phantom(shape, seed=0, device=dev)
    img = (mps * gt.unsqueeze(0)).contiguous()
    ksp = torch.fft.fftshift(torch.fft.fftn(torch.fft.ifftshift(img, dim=(1, 2, 3)), dim=(1, 2, 3), norm="ortho"),
                             dim=(1, 2, 3))
    del img
    mix = rng.standard_normal((C, C)) * 0.1 + np.eye(C)
    noise_h = (mix @ (rng.standard_normal((C, 65536)) + 1j * rng.standard_normal((C, 65536)))).astype(np.complex64)
    noise = torch.as_tensor(noise_h, device=dev)
    sl = tuple(slice(n // 2 - 12, n // 2 + 12) for n in shape)
    calib_h = ksp[(slice(None),) + sl].permute(1, 2, 3, 0).contiguous().cpu().numpy()
    res, nbytes = {}, {}

    def timed(name, fn, reps):
        out = fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = fn()
        torch.cuda.synchronize()
        res[name] = round((time.perf_counter() - t0) * 1e3 / reps, 2)
        log(f"{name}: {res[name]} ms")
        return out

    white, W = timed("noise_whiten", lambda: me.noise_whiten(ksp, noise), max(1, args.steps))
    nbytes["noise_whiten"] = 2 * ksp.numel() * 8 + noise.numel() * 8
    maps = timed("rx_espirit_3d", lambda: me.rx_espirit_3d(calib_h, shape, kernel_size=(5, 5, 5)), 1)
    kept, sweeps = me.rx_espirit_3d.kept, me.rx_espirit_3d.sweeps
    maps_c = maps.permute(3, 0, 1, 2).contiguous()
    del maps
    cc = timed("coil_compression", lambda: me.coil_compression(maps_c, white, perc=0.95), max(1, args.steps))
    ncc = int(cc[0].shape[0])
    nbytes["coil_compression"] = 2 * maps_c.numel() * 8 + white.numel() * 8 + ncc * (maps_c[0].numel() * 8) * 2
    del cc
    gbs = {k: round(nbytes[k] / (res[k] * 1e-3) / 1e9, 1) for k in nbytes}
    out = {
        "metric": "calibration chain ms at 256x256x192, 32 coils (noise whitening, ESPIRiT, coil compression)",
        "value": round(sum(res.values()), 1), "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "higher_is_better": False, "dtype": "c64 data, f64 small matrices", "data": "synthetic",
        "config": {"workload": "ksp 32x256x256x192 complex64, noise 32x65536, ESPIRiT calib 24^3 kernel 5^3 "
                               "eig_thresh 0.02 mask_thresh 0.99, coil compression perc 0.95",
                   "espirit_kept": kept, "jacobi_sweeps": sweeps, "compressed_coils": ncc},
        "ms": res,
        "hbm_gbs": gbs,
    }
    if not args.no_cpu_baseline:
        from oracle import motion_oracle as O
        nthreads = min(16, os.cpu_count() or 1)
        ksp_h = ksp.cpu().numpy()
        mps_h = maps_c.cpu().numpy()
        del ksp, maps_c
        frac = 16  # 1/16 of the voxels for the voxel-wise stages
        nv = ksp_h[0].size
        kv = ksp_h.reshape(C, -1)[:, :nv // frac]
        mv = mps_h.reshape(C, -1)[:, :nv // frac]
        t0 = time.perf_counter()
        O.noise_whiten(kv, noise_h)
        t_w = (time.perf_counter() - t0) * frac
        t0 = time.perf_counter()
        O.coil_compression(mv, kv, perc=0.95)
        t_c = (time.perf_counter() - t0) * frac
        t0 = time.perf_counter()
        O.rx_espirit_3d(calib_h, shape, kernel_size=(5, 5, 5), z_slices=[shape[2] // 2, shape[2] // 4])
        t_e2 = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.rx_espirit_3d(calib_h, shape, kernel_size=(5, 5, 5), z_slices=[])
        t_e0 = time.perf_counter() - t0
        t_e = t_e0 + (t_e2 - t_e0) / 2 * shape[2]
        out["cpu_baseline"] = {"value": round(t_w + t_c + t_e, 1), "unit": "s", "cores": nthreads, "kind": "port",
                               "s": {"noise_whiten": round(t_w, 2), "rx_espirit_3d": round(t_e, 1),
                                     "coil_compression": round(t_c, 2)},
                               "sample": f"FP64 oracle: whitening and compression on 1/{frac} of the voxels, ESPIRiT "
                                         "calibration-matrix SVD in full plus 2 of 192 z-slices of the per-voxel "
                                         "stage, each scaled to the full volume"}
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); shot sharding across GPUs is a design note, "
              f"not part of this build", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    if args.copy_calibration:
        copy_calibration(torch)
        return
    import __graft_entry__ as ge

    me = ge.load_package()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.grad:
        grad_bench(args, torch, me, dev)
        return
    if args.batch:
        batch_bench(args, torch, me, dev)
        return
    if args.calib:
        calib_bench(args, torch, me, dev)
        return
    if args.plain:
        plain_bench(args, torch, me, dev)
        return
    t_setup = time.perf_counter()
    log(f"generating synthetic data for {args.config}")
    data = me.simulation.make_config(args.config, device=dev)
    cfg = data["config"]
    shape, ncoils, nshots = cfg["shape"], cfg["coils"], cfg["nshots"]
    kgrid, rkgrid = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    app = me.JointRecon(data["ksp"], data["mps"], data["A"], kgrid, rkgrid, max_cg_iter=3, max_nm_iter=1,
                        compute_objective=not args.no_objective)
    del data
    tuning = None
    if not args.no_tune:
        tuning = me.batched.tune_problem(app.problem)
        log(f"line-pass tuning (lines per axis, ms per candidate): {tuning}")
    torch.cuda.synchronize()
    log(f"setup done in {time.perf_counter() - t_setup:.1f}s; shot batch {app.problem.batch}")

    objectives = []

    def step():
        # every step also yields one objective value: that of the state it
        # starts from, out of the CG's first encoding of that state
        buf = None if args.no_objective else torch.empty(nshots, dtype=torch.float64, device=dev)
        app.step(prev_objective=buf)
        if buf is not None:
            objectives.append(buf)

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"warmup step {i + 1}/{args.warmup} done")
    pb = app.problem
    timer = KernelTimer(torch)
    orig_encode, orig_decode, orig_rigid = pb._encode, pb._decode, pb._rigid
    pb._encode = timer.wrap("encode_forward", orig_encode)
    pb._decode = timer.wrap("encode_adjoint", orig_decode)
    pb._rigid = timer.wrap("rigid", orig_rigid)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    pb._encode, pb._decode, pb._rigid = orig_encode, orig_decode, orig_rigid
    log(f"timed {args.steps} steps in {t1 - t0:.3f}s")
    ms = (t1 - t0) * 1e3 / args.steps
    # dominant kernel: the gathered forward encoding (k_encode_fwd)
    byts, durs = [], []
    for (s, e, a, k), d in zip(timer.events["encode_forward"], timer.durations("encode_forward")):
        s0, s1 = a[2], a[3]
        byts.append(encode_bytes(pb, s0, s1, k.get("meas", False)))
        durs.append(d)
    achieved = float(np.sum(byts) / np.sum(durs) / 1e9)
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                frac=round(achieved / PEAK_HBM_GBS, 4), traffic=read_traffic(),
                kernel="k_encode_fwd", launches=len(durs), avg_launch_ms=round(1e3 * float(np.mean(durs)), 3),
                bytes_per_launch=float(np.mean(byts)))
    # the adjoint encoding (E^H), the second kernel of the normal operator
    abyts, adurs = [], []
    for (s, e, a, k), d in zip(timer.events.get("encode_adjoint", []), timer.durations("encode_adjoint")):
        abyts.append(decode_bytes(pb, a[1], a[2]))
        adurs.append(d)
    roof_adj = None
    if adurs:
        aach = float(np.sum(abyts) / np.sum(adurs) / 1e9)
        roof_adj = dict(bound="hbm", achieved=round(aach, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                        frac=round(aach / PEAK_HBM_GBS, 4), kernel="k_encode_adj_wave", launches=len(adurs),
                        avg_launch_ms=round(1e3 * float(np.mean(adurs)), 3), bytes_per_launch=float(np.mean(abyts)))
    # the line passes: forward transforms of one image to the shot batch (T p
    # in every normal operator), 10 passes per call
    tbyts, tdurs = [], []
    for (s, e, a, k), d in zip(timer.events.get("rigid", []), timer.durations("rigid")):
        x_bs, theta, mode = a[1], a[3], a[4]
        if (mode == 0 and x_bs == 0 and not k.get("flags", 0) and k.get("x_parts", 1) == 1 and len(a) <= 5
                and theta.shape[0] == pb.batch):  # full shot batches (not the one-image re-registration)
            tbyts.append(transform_bytes(pb, theta.shape[0]))
            tdurs.append(d)
    roof_tr = None
    if tdurs:
        tach = float(np.sum(tbyts) / np.sum(tdurs) / 1e9)
        roof_tr = dict(bound="hbm", achieved=round(tach, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                       frac=round(tach / PEAK_HBM_GBS, 4), kernel="k_linepass (10 passes per transform)",
                       launches=len(tdurs), avg_launch_ms=round(1e3 * float(np.mean(tdurs)), 3),
                       bytes_per_launch=float(np.mean(tbyts)))
    res = {
        "metric": "joint-recon outer iters/s at 256x256x192, 32 coils, 64 shots",
        "value": round(args.steps / (t1 - t0), 4),
        "unit": "outer_iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {shape[0]}x{shape[1]}x{shape[2]} complex64, {ncoils} coils, "
                               f"{nshots} DISORDER shots, 3 CG + 1 LM per outer iteration"
                               + ("" if not args.no_objective else ", objective skipped"),
                   "shape": list(shape), "coils": ncoils, "shots": nshots, "cg_iters": 3,
                   "objective": not args.no_objective, "parallelism": "single GPU",
                   "line_tuning": None if tuning is None else list(tuning[0])},
        "roofline": roof,
        "roofline_adjoint": roof_adj,
        "roofline_transform": roof_tr,
    }
    if not args.no_cpu_baseline:
        log("timing the CPU baseline sample")
        res["cpu_baseline"] = cpu_baseline(tuple(shape), ncoils, nshots, 3)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
