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
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the CPU-baseline leg: the oracle timing and the small-size FP64-oracle accuracy check")
    ap.add_argument("--no-objective", action="store_true", help="skip the per-iteration objective")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuners at setup (line passes and forward encoding; also with --grad, and at "
                         "every level with --full)")
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
    ap.add_argument("--full", action="store_true",
                    help="config 5's whole workload instead of the headline: the 3-level pyramid (x4, x2, x1) with "
                         "--iters joint iterations scaled per level and --final-iters joint iterations at full "
                         "resolution; time to solution, motion and image errors (prints its own JSON line)")
    ap.add_argument("--iters", type=int, default=100, help="n_joint_iters of --full")
    ap.add_argument("--final-iters", type=int, default=100, help="joint outer iterations at full resolution (--full)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (three times), the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE to bytes")
    return ap.parse_args()


COUNTERS_FILE = os.path.join(ROOT, "profiles", "counters_256x256x192_c32_s64.md")


def source_hash():
    """Hash of the library's sources (csrc/ and include/): identifies the
    build a counter profile was taken on."""
    import hashlib

    h = hashlib.sha256()
    for d in (os.path.join(ROOT, "motion-estimation_amd", "csrc"), os.path.join(ROOT, "include")):
        for f in sorted(os.listdir(d)):
            if f.endswith((".hip", ".cpp", ".h")):
                h.update(f.encode())
                with open(os.path.join(d, f), "rb") as fh:
                    h.update(fh.read())
    return h.hexdigest()[:16]


def read_counters():
    """Per-kernel rocprofv3 counter figures of the headline workload (HBM
    bytes from FETCH_SIZE and WRITE_SIZE scaled by a plain copy, VALU
    instructions per element and VALU issue from the SQ counters, kernel-trace
    duration), from the machine-readable block of the committed summary, and
    whether that profile was taken on these sources."""
    try:
        txt = open(COUNTERS_FILE).read()
        block = txt.split("```json", 1)[1].split("```", 1)[0]
        data = json.loads(block)
    except (OSError, IndexError, ValueError):
        return None
    data["same_build"] = data.get("build") == source_hash()
    return data


def counter_fields(counters, kernel, algo_bytes, ms):
    """Counter evidence for one kernel beside its algorithmic roofline."""
    if not counters or kernel not in counters.get("kernels", {}):
        return dict(traffic=None, counters="missing")
    k = counters["kernels"][kernel]
    traffic = k["fetch_bytes"] + k["write_bytes"]
    out = dict(traffic=traffic, hbm_counter_gbs=round(traffic / (k["avg_ms"] * 1e-3) / 1e9, 1),
               hbm_counter_frac=round(traffic / (k["avg_ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
               traffic_over_algorithmic=round(traffic / algo_bytes, 3), profiled_avg_ms=k["avg_ms"],
               profile=os.path.relpath(COUNTERS_FILE, ROOT), profile_same_build=counters["same_build"])
    if "valu_issue" in k:
        out.update(valu_issue=k["valu_issue"], valu_per_element=k["valu_per_elem"])
    return out


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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


def host_threads():
    """(threads used, cores in the affinity set): every core of the affinity
    set, bounded by the CPU share the host grants this process when it states
    one in OMP_NUM_THREADS."""
    aff = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS", "")
    n = min(aff, int(share)) if share.isdigit() and int(share) > 0 else aff
    return n, aff


def cpu_baseline(shape, ncoils, nshots, n_cg):
    """Times the oracle's complex64 restatement (dense masks, coil-batched
    3-D FFTs, dense dots, as the reference runs) on one shot of each unit of
    an outer iteration at the full shape, and composes the outer iteration by
    the reference's counts (joint_recon.py:92-102): per shot one right-hand
    side term, n_cg + 1 normal-operator terms, one LM update
    (transform_estimation.py:39-144) and one objective term, plus one
    transform for the mean re-registration."""
    import numpy as np

    from oracle import motion_oracle as O

    threads, aff = host_threads()
    O.set_workers(threads)
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    mps = np.full((ncoils, *shape), 1 / np.sqrt(ncoils), dtype=np.complex64)
    ksp = (mps * np.roll(x, 3, axis=0)[None]).astype(np.complex64)  # any data of the right shape
    mask = np.zeros((1, shape[0], shape[1], 1), dtype=bool)  # one shot: 512 of the plane's lines
    mask[0, ::8, ::16] = True
    M = np.ones(shape, dtype=bool)
    kg, rkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    th = np.array([[0.5, -0.3, 0.2, 0.02, -0.01, 0.015]])
    t = {}

    def timed(name, fn):
        t0 = time.perf_counter()
        fn()
        t[name] = time.perf_counter() - t0

    timed("T", lambda: O.rigid_apply(x, th[0], kg, rkg, prec="c64"))
    timed("rhs_shot", lambda: O.cg_rhs(ksp, mps, mask, th, kg, rkg, M, prec="c64"))
    timed("normal_shot", lambda: O.normal_op(x, mps, mask, th, kg, rkg, M, prec="c64"))
    timed("objective_shot", lambda: O._vdot_real(*(2 * [O.shot_residual(ksp, mps, mask[0], x, th[0], kg, rkg,
                                                                             prec="c64")]), "c64"))
    timed("lm_shot", lambda: O.estimate_transform(ksp, mps, mask, x, kg, rkg, np.ones(1), np.zeros(1, bool), th.copy(),
                                                  prec="c64"))
    lm = t["lm_shot"] - t["T"]  # its mean re-registration transform is counted once below
    outer = nshots * (t["rhs_shot"] + (n_cg + 1) * t["normal_shot"] + lm + t["objective_shot"]) + t["T"]
    sample = sum(t.values())
    return dict(value=1.0 / outer, unit="outer_iters/s", cores=threads, kind="port",
                cpu_model=cpu_model(), affinity_cores=aff,
                sample=(f"oracle complex64 restatement at {shape[0]}x{shape[1]}x{shape[2]}, {ncoils} coils, one shot of "
                        f"each unit on {threads} threads: right-hand side {t['rhs_shot']:.1f}s, normal operator "
                        f"{t['normal_shot']:.1f}s, LM update (residual, 6 Jacobian columns, 27 dense dots, candidate) "
                        f"{lm:.1f}s, objective {t['objective_shot']:.1f}s, transform {t['T']:.2f}s = {sample:.0f}s; "
                        f"composed into one outer iteration ({nshots} shots, {n_cg} CG) = {outer:.0f}s"),
                unit_seconds={k: round(v, 2) for k, v in t.items()})


def small_problem(me, shape, nc, seed, rot_deg=3.0, trans_vox=1.5):
    """A seeded motion-corrupted problem small enough for the FP64 oracle
    (centred frame, numpy): phantom, coil maps, DISORDER shots (tile 4 x 4,
    acceleration 2), uniform random motion, k-space corrupted by the oracle's
    generate_corrupted_kspace plus noise at 1e-3 of the peak."""
    import numpy as np

    from oracle import motion_oracle as O

    sim = me.simulation
    gt = sim.phantom(shape, seed=seed).numpy().astype(np.complex128)
    mps = sim.coil_maps(shape, nc, seed=seed + 1).numpy().astype(np.complex128)
    A, _ = sim.disorder_sampling(shape, (4, 4), 2)
    t_true = sim.motion_uniform(A.shape[0], rot_deg=rot_deg, trans_vox=trans_vox, seed=seed + 2)
    ksp = O.generate_corrupted_kspace(gt, mps, A, t_true)
    rng = np.random.default_rng(seed + 3)
    ksp = ksp + 1e-3 * np.abs(ksp).max() * (rng.standard_normal(ksp.shape) + 1j * rng.standard_normal(ksp.shape))
    return gt, mps, A, ksp


def _motion_err(t, t_ref):
    import numpy as np

    d = np.abs(np.asarray(t) - np.asarray(t_ref))
    return float(max(d[:, :3].max(), np.degrees(d[:, 3:]).max()))


def _rel(a, b):
    import numpy as np

    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def oracle_accuracy(me, dev):
    """Error of the measured path against the FP64 oracle on the same seeded
    inputs at a size the oracle finishes in about a second: three outer
    iterations of JointRecon (right-hand side, 3 CG steps, LM, objective) at
    32 x 24 x 16 with 4 coils and 8 DISORDER shots, motion-corrupted data with
    noise."""
    import numpy as np

    from oracle import motion_oracle as O

    shape = (32, 24, 16)
    gt, mps, A, ksp = small_problem(me, shape, 4, 0)
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    okg, orkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    app = me.JointRecon(ksp, mps, A, kg, rkg, max_cg_iter=3, max_joint_iter=3)
    x, t = app.run()
    ref = O.JointRecon(ksp, mps, A, okg, orkg, max_cg_iter=3, max_joint_iter=3)
    xr, tr = ref.run()
    x, t = x.cpu().numpy(), t.cpu().numpy()
    d = np.abs(t - tr)
    return dict(workload="JointRecon, 3 outer iterations of 3 CG + LM, 32x24x16, 4 coils, 8 shots",
                image_rel_l2=float(np.linalg.norm(x - xr) / np.linalg.norm(xr)),
                motion_max_abs_deg_vox=float(max(d[:, :3].max(), np.degrees(d[:, 3:]).max())),
                objective_rel=float(np.max(np.abs(app.objective_history - ref.objective_history)
                                           / ref.objective_history)),
                damping_identical=bool(np.array_equal(app.damp.cpu().numpy(), ref.damp)),
                bars=dict(image_rel_l2=1e-4, motion=1e-3, objective_rel=1e-4))


def grad_accuracy(me, dev):
    """The data-consistency loss and both gradients (image, motion) of the
    measured path against the FP64 oracle's dc_loss / dc_loss_grad at
    16 x 12 x 8, 4 coils, 24 DISORDER shots, random data and motion."""
    import numpy as np
    import torch

    from oracle import motion_oracle as O

    shape, nc = (16, 12, 8), 4
    sim = me.simulation
    rng = np.random.default_rng(40)
    mps = np.fft.ifftshift(sim.coil_maps(shape, nc, seed=41).numpy().astype(np.complex128), axes=(-3, -2, -1))
    A, _ = sim.disorder_sampling(shape, (4, 4), 2)
    A = np.fft.ifftshift(A, axes=(1, 2))
    ksp = rng.standard_normal((nc, *shape)) + 1j * rng.standard_normal((nc, *shape))
    t = sim.motion_uniform(A.shape[0], rot_deg=4.0, trans_vox=1.5, seed=42)
    x0 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    okg, orkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    pb = me.encoding.SenseProblem(ksp, mps, A, kg, rkg)
    x = torch.as_tensor(x0.astype(np.complex64), device=dev).requires_grad_(True)
    th = torch.as_tensor(t, device=dev).requires_grad_(True)
    loss = me.data_consistency_loss(x, th, pb)
    loss.backward()
    ref = O.dc_loss(x0, t, ksp, mps, A, okg, orkg)
    gx, gt = O.dc_loss_grad(x0, t, ksp, mps, A, okg, orkg)
    return dict(workload=f"data-consistency loss and gradients, 16x12x8, 4 coils, {A.shape[0]} shots",
                loss_rel=float(abs(float(loss.detach()) - ref) / ref),
                grad_image_rel_l2=_rel(x.grad.cpu().numpy(), gx), grad_motion_rel_l2=_rel(th.grad.cpu().numpy(), gt),
                bars=dict(loss_rel=1e-4, grad_image_rel_l2=1e-4, grad_motion_rel_l2=1e-4))


def normal_accuracy(me, dev):
    """The forward encoding E_s T_s of every shot and the normal operator
    sum_s T_s^H E_s^H E_s T_s of the measured path against the FP64 oracle's
    encode_forward / normal_op at 32 x 24 x 16, 4 coils, 8 DISORDER shots,
    random image and motion (uncentred frame, as the reference's operators)."""
    import numpy as np
    import torch

    from oracle import motion_oracle as O

    shape, nc = (32, 24, 16), 4
    sim = me.simulation
    mps = np.fft.ifftshift(sim.coil_maps(shape, nc, seed=51).numpy().astype(np.complex128), axes=(-3, -2, -1))
    A, _ = sim.disorder_sampling(shape, (4, 4), 2)
    A = np.fft.ifftshift(A, axes=(1, 2))
    t = sim.motion_uniform(A.shape[0], rot_deg=3.0, trans_vox=1.5, seed=52)
    rng = np.random.default_rng(53)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    okg, orkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
    pb = me.encoding.SenseProblem(np.zeros((nc, *shape), np.complex64), mps, A, kg, rkg)
    pb.theta.copy_(torch.as_tensor(t))
    Y = pb.forward(pb.image_to_plane(x)).cpu().numpy()
    pl, n2 = pb.plan, shape[2]
    fwd = []
    for s in range(A.shape[0]):
        hyb = np.fft.ifft(O.encode_forward(x, mps, A[s], t[s], okg, orkg), axis=-1, norm="ortho")
        o, c = int(pl.off_h[s]), int(pl.cnt_h[s])
        ref = hyb[:, pl.k0_h[o:o + c], pl.k1_h[o:o + c], :]
        got = np.transpose(Y[o * n2 * nc:(o + c) * n2 * nc].reshape(n2, nc, c), (1, 2, 0))
        fwd.append(_rel(got, ref))
    out = pb.zeros_image()
    pb.normal(pb.image_to_plane(x), out)
    M = np.sum(np.abs(mps) ** 2, axis=0) > 0.1
    ref = O.normal_op(x, mps, A, t, okg, orkg, M)
    return dict(workload=f"E_s T_s per shot and the normal operator, 32x24x16, 4 coils, {A.shape[0]} shots",
                forward_rel_l2_max_over_shots=float(max(fwd)),
                normal_rel_l2=_rel(pb.image_from_plane(out).cpu().numpy(), ref),
                bars=dict(forward_rel_l2_max_over_shots=1e-4, normal_rel_l2=1e-4))


def batch_accuracy(me, dev):
    """Two volumes of different size batched (graph replay over 2 streams,
    three outer iterations of 3 CG + LM) against the FP64 oracle's JointRecon
    of each volume alone."""
    from oracle import motion_oracle as O

    probs = [((32, 24, 16), 4, 60), ((24, 24, 12), 3, 70)]
    vols, data = [], []
    for shape, nc, seed in probs:
        _, mps, A, ksp = small_problem(me, shape, nc, seed)
        kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
        vols.append(dict(ksp=ksp, mps=mps, shot_mask=A, kgrid=kg, rkgrid=rkg))
        data.append((shape, mps, A, ksp))
    bj = me.BatchedJointRecon(vols, nstreams=2)
    try:
        imgs, ts = bj.run(3)
        got = [(i.cpu().numpy(), t.cpu().numpy()) for i, t in zip(imgs, ts)]
    finally:
        bj.close()
    img_err, mot_err = [], []
    for (shape, mps, A, ksp), (xg, tg) in zip(data, got):
        okg, orkg = O.compute_transform_grids_voxel(shape, (1, 1, 1))
        xr, tr = O.JointRecon(ksp, mps, A, okg, orkg, max_cg_iter=3, max_joint_iter=3, compute_objective=False).run()
        img_err.append(_rel(xg, xr))
        mot_err.append(_motion_err(tg, tr))
    return dict(workload="BatchedJointRecon graph replay, 3 outer iterations of 3 CG + LM, volumes 32x24x16 "
                         f"(4 coils, {data[0][2].shape[0]} shots) and 24x24x12 (3 coils, {data[1][2].shape[0]} shots), "
                         "each against the oracle alone",
                image_rel_l2_max=float(max(img_err)), motion_max_abs_deg_vox=float(max(mot_err)),
                bars=dict(image_rel_l2_max=1e-4, motion_max_abs_deg_vox=1e-3))


def pyramid_accuracy(me, dev):
    """The pyramid driver (one coarse JointRecon level, then CG-only at full
    resolution) against the FP64 oracle's pyramid_reconstruction at
    32 x 32 x 8, 4 coils, DISORDER shots."""
    from oracle import motion_oracle as O

    _, mps, A, ksp = small_problem(me, (32, 32, 8), 4, 80, rot_deg=2.0, trans_vox=1.0)
    kw = dict(n_spatial_levels=2, n_joint_iters=1, final_cg_iter=20)
    x, t = me.pyramid_reconstruction(ksp, mps, A, A.shape[0], compute_objective=False, **kw)
    xr, tr = O.pyramid_reconstruction(ksp, mps, A, A.shape[0], **kw)
    return dict(workload=f"pyramid, 2 levels (x2 JointRecon 2 outer, x1 CG 20), 32x32x8, 4 coils, {A.shape[0]} shots",
                image_rel_l2=_rel(x, xr), motion_max_abs_deg_vox=_motion_err(t, tr),
                bars=dict(image_rel_l2=1e-4, motion_max_abs_deg_vox=1e-3))


def calib_accuracy(me, dev):
    """The three calibration stages against the FP64 oracle at small sizes:
    noise whitening (12 coils, 16 x 12 x 8), coil compression (16 coils,
    24 x 20 x 12, perc 0.95: singular values and the compression subspace)
    and ESPIRiT (4 coils, 24 x 20 x 16, calibration 12 x 12 x 10, kernel 3^3:
    kept vectors, singular values and the maps where both are non-zero)."""
    import numpy as np

    from oracle import motion_oracle as O

    sim = me.simulation
    rng = np.random.default_rng(90)
    C = 12
    mix = rng.standard_normal((C, C)) + 1j * rng.standard_normal((C, C))
    noise = (mix @ (rng.standard_normal((C, 5000)) + 1j * rng.standard_normal((C, 5000)))).astype(np.complex64)
    ksp = (rng.standard_normal((C, 16, 12, 8)) + 1j * rng.standard_normal((C, 16, 12, 8))).astype(np.complex64)
    wg, _ = me.noise_whiten(ksp, noise)
    wr, _ = O.noise_whiten(ksp, noise)
    shape = (24, 20, 12)
    mps = sim.coil_maps(shape, 16, seed=91).numpy()
    kc = (rng.standard_normal((16, *shape)) + 1j * rng.standard_normal((16, *shape))).astype(np.complex64)
    me.coil_compression(mps, kc, perc=0.95)
    _, _, Ar, S = O.coil_compression(mps, kc, perc=0.95)
    Ag = me.coil_compression.matrix.cpu().numpy().astype(np.complex128)
    lam = me.coil_compression.singular_values
    imsize, nc = (24, 20, 16), 4
    m4 = sim.coil_maps(imsize, nc, seed=92).numpy().astype(np.complex128)
    img = m4 * sim.phantom(imsize, seed=93).numpy()[None]
    k4 = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(img, axes=(1, 2, 3)), axes=(1, 2, 3), norm="ortho"), axes=(1, 2, 3))
    sl = tuple(slice(n // 2 - c // 2, n // 2 - c // 2 + c) for n, c in zip(imsize, (12, 12, 10)))
    calib = np.ascontiguousarray(np.moveaxis(k4[(slice(None),) + sl], 0, -1)).astype(np.complex64)
    info = {}
    er = O.rx_espirit_3d(calib, imsize, kernel_size=(3, 3, 3), info=info)
    eg = me.rx_espirit_3d(calib, imsize, kernel_size=(3, 3, 3)).cpu().numpy().astype(np.complex128)
    both = np.any(er != 0, axis=-1) & np.any(eg != 0, axis=-1)
    Se, se = info["singular_values"], me.rx_espirit_3d.singular_values
    k = int(np.sum(Se > 1e-3 * Se[0]))
    return dict(workload="noise whitening 12 coils 16x12x8; coil compression 16 coils 24x20x12 perc 0.95; "
                         "ESPIRiT 4 coils 24x20x16, calibration 12x12x10, kernel 3^3",
                whitened_rel_l2=_rel(wg.cpu().numpy(), wr),
                compression_singular_values_max_err_over_s0=float(np.max(np.abs(lam - S)) / S[0]),
                compression_subspace_rel=_rel(Ag.conj().T @ Ag, Ar.conj().T @ Ar),
                espirit_kept_equal=bool(me.rx_espirit_3d.kept == info["kept"]),
                espirit_singular_values_max_err_over_s0=float(np.max(np.abs(se[:k] - Se[:k])) / Se[0]),
                espirit_maps_rel_l2_common_support=_rel(eg[both], er[both]),
                bars=dict(whitened_rel_l2=1e-5, compression_singular_values_max_err_over_s0=1e-6,
                          compression_subspace_rel=1e-6, espirit_singular_values_max_err_over_s0=1e-6,
                          espirit_maps_rel_l2_common_support=1e-4))


def grad_bench(args, torch, me, dev):
    """Forward+backward of the data-consistency loss at c3 with recomputed or
    stored sheared intermediates, plus a central-difference check."""
    import numpy as np

    data = me.simulation.make_config("c3", device=dev)
    cfg = data["config"]
    shape, ncoils, nshots = cfg["shape"], cfg["coils"], cfg["nshots"]
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    pb = me.encoding.SenseProblem(data["ksp"], data["mps"], data["A"], kg, rkg, centered=True)
    tuning = None
    if not args.no_tune:
        tuning = dict(line_passes=me.batched.tune_problem(pb), encoders=me.batched.tune_encoders(pb))
        log(f"grad: tuning {tuning}")
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
                               f"DISORDER shots, gradients w.r.t. image and 48x6 motion",
                   "tuning": None if tuning is None else {
                       "line_passes": tuning["line_passes"][0],
                       "encoders": {str(k): v[0] for k, v in tuning["encoders"].items()}}},
        "ms": {k: round(v, 3) for k, v in res.items()},
        "fd_check": {"directions": 64, "step": h, "median_rel": float(np.median(rel)), "max_rel": float(max(rel)),
                     "max_rel_to_gradient_scale": float(max(scaled))},
    }
    if not args.no_cpu_baseline:
        log("checking the same path against the FP64 oracle at a small size")
        out["accuracy_vs_fp64_oracle"] = grad_accuracy(me, dev)
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
    del pb, mps_u, masks, M, xp, outp
    torch.cuda.empty_cache()
    out["outer_iteration"] = plain_outer(args, torch, me, dev)
    out["speedup_over_plain"]["outer_iteration"] = out["outer_iteration"]["speedup"]
    if not args.no_cpu_baseline:
        log("checking the same path against the FP64 oracle at a small size")
        out["accuracy_vs_fp64_oracle"] = normal_accuracy(me, dev)
    print(json.dumps(out), flush=True)


def plain_outer(args, torch, me, dev):
    """One outer iteration of the joint reconstruction at 256x256x192, 32
    coils, 64 shots (RHS, 3 CG, LM of 64 x 6, objective) through the library
    and through the plain torch.fft restatement of the reference
    (PlainJointRecon), on the same synthetic data; both start from x = 0,
    t = 0, and their first iterations are compared."""
    import numpy as np

    data = me.simulation.make_config("c5", device=dev)
    shape = tuple(data["config"]["shape"])
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    app = me.JointRecon(data["ksp"], data["mps"], data["A"], kg, rkg, max_cg_iter=3, max_nm_iter=1)
    tuning = me.batched.tune_problem(app.problem)
    buf = torch.empty(64, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    app.step()
    torch.cuda.synchronize()
    first_ms = (time.perf_counter() - t0) * 1e3
    x1, t1 = app.image().cpu(), app.t.cpu().numpy()
    steps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        app.step(prev_objective=buf)
    torch.cuda.synchronize()
    ours_ms = (time.perf_counter() - t0) * 1e3 / steps
    log(f"library outer iteration: {ours_ms:.1f} ms (first {first_ms:.1f} ms)")
    del app
    torch.cuda.empty_cache()
    tkg, trkg = me.compute_transform_grids_voxel(shape, (1, 1, 1), xp=torch)
    pj = PlainJointRecon(torch, data["ksp"], data["mps"], data["A"], tkg, trkg, max_cg_iter=3)
    del data
    pj.normal(pj.x)  # FFT plans and allocator warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pj.step()
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3
    log(f"plain outer iteration: {plain_ms:.1f} ms")
    xr = pj.image().cpu()
    d = np.abs(t1 - pj.t)
    return dict(ours_ms=round(ours_ms, 1), plain_ms=round(plain_ms, 1), speedup=round(plain_ms / ours_ms, 2),
                line_tuning=list(tuning[0]),
                first_iteration_image_rel_l2=float(torch.linalg.vector_norm(x1 - xr) / torch.linalg.vector_norm(xr)),
                first_iteration_motion_max_abs=float(max(d[:, :3].max(), np.degrees(d[:, 3:]).max())),
                plain_lm_accepted=int(sum(pj.accepted[0])))


def full_bench(args, torch, me, dev):
    """Config 5's whole workload: a 3-level pyramid (x4, x2, x1) over the
    256x256x192, 32-coil, 64-shot problem, n_joint_iters = 100 (400 and 200
    outer iterations at the coarse levels, pyramid.py:73) and, as the
    configuration asks, 100 joint outer iterations at full resolution
    (final_joint_iters; the reference's final level is CG only,
    pyramid.py:96-115).  Reports time to solution per level, the motion
    error against the per-shot truth and the image error, with the
    uncorrected SENSE image's error beside it."""
    import numpy as np

    data = me.simulation.make_config("c5", device=dev)
    cfg = data["config"]
    shape, nshots = tuple(cfg["shape"]), cfg["nshots"]
    gt = data["gt"]
    # per-shot truth: the mean of the corruption states a shot's samples come
    # from (128 states, two per shot), without its mean over shots (the
    # reconstruction's mean re-registration moves that into the image)
    tc = np.asarray(data["t_corr"])
    per = tc.shape[0] // nshots
    t_true = tc[:per * nshots].reshape(nshots, per, 6).mean(1)
    t_mean = t_true.mean(0)
    iters = args.iters
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, t_est = me.pyramid_reconstruction(data["ksp"], data["mps"], data["A"], nshots, n_spatial_levels=3,
                                         n_joint_iters=iters, final_joint_iters=args.final_iters,
                                         compute_objective=False, progress=log, tune=not args.no_tune)
    total_s = time.perf_counter() - t0
    levels = me.pyramid.pyramid_reconstruction.levels
    log(f"pyramid done in {total_s:.1f}s: " + ", ".join(f"x{l['factor']} {l['seconds']:.1f}s" for l in levels))
    dt = t_est - (t_true - t_mean)
    # image in the mean-motion frame the reconstruction lands in
    kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1))
    T = me.RigidTransformCudaOptimzied(t_mean, kg, rkg)
    ref = torch.fft.fftshift(T.apply(torch.fft.ifftshift(gt.to(torch.complex64))))
    xg = torch.as_tensor(x, device=dev)
    t0 = time.perf_counter()
    # simulation.py:462-465: P = 1/(sum |S|^2 + 1e-6), tol 1e-12
    p_sense = 1 / ((data["mps"].abs() ** 2).sum(0) + 1e-6)
    unc = me.sense_recon(data["ksp"], data["mps"], P=p_sense, tol=1e-12, max_iter=100)
    sense_s = time.perf_counter() - t0

    def err(a):
        return float(torch.linalg.vector_norm(a.abs() - ref.abs()) / torch.linalg.vector_norm(ref.abs()))

    out = {
        "metric": "time to solution of the 3-level pyramid at 256x256x192, 32 coils, 64 shots",
        "value": round(total_s, 1), "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": f"c5: {shape[0]}x{shape[1]}x{shape[2]}, {cfg['coils']} coils, {nshots} shots, "
                               f"128 corruption states of smooth cubic motion (+-5 deg, +-5 voxels), 30 dB noise; "
                               f"levels x4 ({4 * iters} outer), x2 ({2 * iters} outer), x1 "
                               + (f"({args.final_iters} joint outer)" if args.final_iters else "(CG only, 100 steps)"),
                   "n_joint_iters": iters, "final_joint_iters": args.final_iters},
        "levels": [dict(factor=l["factor"], shape=list(l["shape"]), seconds=round(l["seconds"], 2)) for l in levels],
        "motion_error_vs_truth": {"rms_vox": float(np.sqrt(np.mean(dt[:, :3] ** 2))),
                                  "max_vox": float(np.abs(dt[:, :3]).max()),
                                  "rms_deg": float(np.degrees(np.sqrt(np.mean(dt[:, 3:] ** 2)))),
                                  "max_deg": float(np.degrees(np.abs(dt[:, 3:]).max())),
                                  "truth_rms_vox": float(np.sqrt(np.mean((t_true - t_mean)[:, :3] ** 2))),
                                  "truth_rms_deg": float(np.degrees(np.sqrt(np.mean((t_true - t_mean)[:, 3:] ** 2))))},
        "image_error_magnitude_rel_l2": {"corrected": err(xg), "uncorrected_sense": err(unc),
                                         "sense_seconds": round(sense_s, 2)},
    }
    if not args.no_cpu_baseline:
        log("checking the same path against the FP64 oracle at a small size")
        out["accuracy_vs_fp64_oracle"] = pyramid_accuracy(me, dev)
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
    lib = me._lib.load()
    lib.me_line_tuning_clear()
    lib.me_encode_tuning_clear()
    bj.step_eager()  # allocates every workspace
    torch.cuda.synchronize()
    log("warm-up outer iteration done")
    nvol = len(bj.recons)
    res = {}

    def timed(name, fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        # host monotonic clock (the clock of rocprofv3 kernel timestamps): lets a
        # kernel trace of this run attribute dispatches to the timed legs
        log(f"leg {name} starts at monotonic_ns {time.monotonic_ns()}")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        log(f"leg {name} ends at monotonic_ns {time.monotonic_ns()}")
        res[name] = dict(ms_per_iteration=round(dt * 1e3 / args.steps, 2),
                         volume_iters_per_s=round(nvol * args.steps / dt, 2))
        log(f"{name}: {res[name]}")

    timed("one_at_a_time", bj.step_eager)
    timed("graph_default", bj.step)
    bj.close()
    tuning = bj.tune(reps=5)
    log(f"tuning: {tuning}; encoders: {bj.encoder_tuning}")
    timed("one_at_a_time_tuned", bj.step_eager)
    timed("graph_tuned", bj.step)
    bj.close()
    lib.me_line_tuning_clear()
    lib.me_encode_tuning_clear()
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
        "speedup_graph_vs_one_at_a_time": round(res["graph_tuned"]["volume_iters_per_s"]
                                                / res["one_at_a_time_tuned"]["volume_iters_per_s"], 3),
        "speedup_tuned_vs_default": round(res["graph_tuned"]["volume_iters_per_s"]
                                          / res["graph_default"]["volume_iters_per_s"], 3),
        "speedup_vs_untuned_one_at_a_time": round(res["graph_tuned"]["volume_iters_per_s"]
                                                  / res["one_at_a_time"]["volume_iters_per_s"], 3),
        "tuning": {"x".join(map(str, k)): dict(lines=list(v[0]), ms=list(v[1])) for k, v in tuning.items()},
        "encoder_tuning": {f"{'x'.join(map(str, k[0]))} c{k[1]} s{k[2]}": {str(m): dict(variant=v[0], ms=list(v[1]))
                                                                         for m, v in e.items()}
                           for k, e in bj.encoder_tuning.items()},
    }
    if not args.no_cpu_baseline:
        log("checking the same path against the FP64 oracle at a small size")
        out["accuracy_vs_fp64_oracle"] = batch_accuracy(me, dev)
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
        nthreads, aff = host_threads()
        O.set_workers(nthreads)
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
                               "cpu_model": cpu_model(), "affinity_cores": aff,
                               "s": {"noise_whiten": round(t_w, 2), "rx_espirit_3d": round(t_e, 1),
                                     "coil_compression": round(t_c, 2)},
                               "sample": f"FP64 oracle: whitening and compression on 1/{frac} of the voxels, ESPIRiT "
                                         "calibration-matrix SVD in full plus 2 of 192 z-slices of the per-voxel "
                                         "stage, each scaled to the full volume"}
        out["accuracy_vs_fp64_oracle"] = calib_accuracy(me, dev)
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); shot sharding across GPUs is designed but "
              f"not implemented", file=sys.stderr)
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
    if args.full:
        full_bench(args, torch, me, dev)
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
    enc_tuning = None
    if not args.no_tune:
        tuning = me.batched.tune_problem(app.problem)
        log(f"line-pass tuning (lines per axis, ms per candidate): {tuning}")
        enc_tuning = me.batched.tune_encoders(app.problem)
        log(f"forward-encoding tuning (max samples per shot: variant, ms per variant): {enc_tuning}")
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
    counters = read_counters()
    # dominant kernel: the gathered forward encoding (k_encode_fwd)
    byts, durs = [], []
    for (s, e, a, k), d in zip(timer.events["encode_forward"], timer.durations("encode_forward")):
        s0, s1 = a[2], a[3]
        byts.append(encode_bytes(pb, s0, s1, k.get("meas", False)))
        durs.append(d)
    achieved = float(np.sum(byts) / np.sum(durs) / 1e9)
    enc_model = ("per shot V + C*V + 8*C*K_s: the image and every coil map read once per shot, the gathered samples "
                 "written; the single-pass model with rho = 0, since the fused kernel keeps the coil intermediate "
                 "in LDS and registers (no HBM round trip)")
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                frac=round(achieved / PEAK_HBM_GBS, 4), kernel="k_encode_fwd", launches=len(durs),
                avg_launch_ms=round(1e3 * float(np.mean(durs)), 3), bytes_per_launch=float(np.mean(byts)),
                model=enc_model, limiter="VALU issue (not HBM): the maps are re-read per shot from L2, so the "
                                         "algorithmic bytes exceed the HBM traffic")
    roof.update(counter_fields(counters, "k_encode_fwd", float(np.mean(byts)), roof["avg_launch_ms"]))
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
                        avg_launch_ms=round(1e3 * float(np.mean(adurs)), 3), bytes_per_launch=float(np.mean(abyts)),
                        model=enc_model.replace("written", "read, one image written"),
                        limiter="VALU issue and LDS latency (not HBM)")
        roof_adj.update(counter_fields(counters, "k_encode_adj_wave", float(np.mean(abyts)),
                                       roof_adj["avg_launch_ms"]))
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
                       bytes_per_launch=float(np.mean(tbyts)),
                       model="each of the 10 single-axis passes reads and writes every shot's volume once "
                             "(the first reads the one input image)", limiter="HBM")
        roof_tr.update(counter_fields(counters, "transform_10_passes", float(np.mean(tbyts)),
                                      roof_tr["avg_launch_ms"]))
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
                   "line_tuning": None if tuning is None else list(tuning[0]),
                   "encoder_tuning": None if enc_tuning is None else {str(k): v[0] for k, v in enc_tuning.items()}},
        "roofline": roof,
        "roofline_adjoint": roof_adj,
        "roofline_transform": roof_tr,
    }
    if not args.no_cpu_baseline:
        log("checking the same path against the FP64 oracle at a small size")
        res["accuracy_vs_fp64_oracle"] = oracle_accuracy(me, dev)
        log("timing the CPU baseline sample")
        res["cpu_baseline"] = cpu_baseline(tuple(shape), ncoils, nshots, 3)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
