"""Many independent joint reconstructions advanced together.

Each volume keeps its own :class:`JointRecon` state (matrix size, coil count,
shot count and DISORDER pattern may all differ).  ``step()`` advances every
volume by one outer iteration (joint_recon.py:92-100: CG, then the LM update
and x *= M).  The first call runs eagerly (it allocates the per-volume
workspaces and the E^H y cache); after that one outer iteration of all
volumes is recorded once into a HIP graph and replayed:

* the host issues one graph launch per iteration instead of a few hundred
  kernel launches per volume;
* the volumes are dealt round-robin to ``nstreams`` side streams joined to the
  capturing stream through events, so the graph holds independent branches
  that the GPU runs concurrently, filling the device when single volumes are
  too small to do so.

``tune()`` runs the autotuners on each distinct volume geometry before
capture (the line passes' lines per workgroup, the forward encoding's coils
per workgroup and samples per thread), so the recorded kernels use the
fastest choices.  Replays compute exactly what the eager calls compute: the kernels
and their per-volume order are the same.  No counterpart in the reference,
which reconstructs one volume per process.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .joint_recon import JointRecon


def tune_problem(pb, reps=5):
    """Run the line-pass autotuner for one problem's geometry (plane-major
    layout, its shot batch) on its scratch buffer; returns (chosen lines per
    axis, mean ms per candidate).  Results do not depend on the choice."""
    n0, n1, n2 = pb.shape
    U = pb._ubuf()
    chosen = (C.c_int * 3)()
    times = (C.c_float * 9)()
    _lib.call("me_tune_line_passes", _lib.ptr(U), _lib.ptr(U), pb.vol, int(pb.batch), _lib.int3(pb.shape),
              _lib.ll3((n1, 1, n0 * n1)), int(reps), chosen, times, _lib.stream(pb.device))
    return tuple(chosen), tuple(round(t, 4) for t in times)


def tune_encoders(pb, reps=3):
    """Pick the forward-encoding variant (coils per workgroup, samples per
    thread) for each shot batch of one problem by timing all of them on its
    buffers; returns {max_cnt: (chosen variant, mean ms per variant)}.
    The samples do not depend on the choice."""
    out = {}
    pl = pb.plan
    U, Y = pb._ubuf(), pb._ybuf()
    for s0, s1 in pb.batches():
        maxc = int(pl.cnt_h[s0:s1].max()) if s1 > s0 else 0
        if maxc in out:
            continue
        chosen = C.c_int()
        times = (C.c_float * 4)()
        _lib.call("me_tune_encode_forward", _lib.ptr(U), pb.vol, _lib.ptr(pb.maps), pb.ncoils, s1 - s0,
                  _lib.int3(pb.shape), _lib.ptr(pl.sk0), _lib.ptr(pl.sk1), _lib.ptr(pl.off[s0:s1]),
                  _lib.ptr(pl.cnt[s0:s1]), maxc, _lib.ptr(Y), int(reps), C.byref(chosen), times,
                  _lib.stream(pb.device))
        out[maxc] = (int(chosen.value), tuple(round(t, 4) for t in times))
    return out


class BatchedJointRecon:
    def __init__(self, volumes, max_cg_iter=3, max_nm_iter=1, nstreams=4):
        """volumes: iterable of dicts with keys ksp, mps, shot_mask, kgrid,
        rkgrid and optionally t0, as for :class:`JointRecon` (centred frame)."""
        self.device = _lib.require_gpu()
        self.recons = [JointRecon(v["ksp"], v["mps"], v["shot_mask"], v["kgrid"], v["rkgrid"], t0=v.get("t0"),
                                  max_cg_iter=max_cg_iter, max_nm_iter=max_nm_iter, max_joint_iter=0,
                                  compute_objective=False, cg_check_every=0) for v in volumes]
        if not self.recons:
            raise ValueError("no volumes")
        # default-priority side streams
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(max(1, int(nstreams)))]
        self._events = None
        self._exec = None
        self._capture_stream = None
        self.steps_done = 0
        self.tuning = {}
        self.encoder_tuning = {}

    # ------------------------------------------------------------ tuning
    def tune(self, reps=5):
        """Autotune every distinct volume geometry before capture: the plain
        line passes' lines per workgroup (plane-major layout, the problem's
        shot batch) and the forward encoding's variant (coils per workgroup,
        samples per thread) for each (shape, coils, largest shot).  Returns
        {shape: (chosen lines per axis, per-candidate ms)} and keeps the
        encoder picks in ``encoder_tuning``."""
        done, enc = {}, {}
        for r in self.recons:
            pb = r.problem
            if pb.shape not in done:
                done[pb.shape] = tune_problem(pb, reps)
            key = (pb.shape, pb.ncoils, pb.plan.max_cnt)
            if key not in enc:
                enc[key] = tune_encoders(pb, max(1, reps // 2))
        self.tuning = done
        self.encoder_tuning = enc
        return done

    # ------------------------------------------------------------ stepping
    def step_eager(self):
        """One outer iteration of every volume, one volume after another on
        the current stream (the one-at-a-time path)."""
        for r in self.recons:
            r.step()
        self.steps_done += 1

    def _capture(self):
        # capture needs a stream of its own (not the legacy default stream)
        s0 = torch.cuda.Stream(device=self.device)
        s0.wait_stream(torch.cuda.current_stream(self.device))
        k = len(self.streams)
        fork = torch.cuda.Event()
        joins = [torch.cuda.Event() for _ in range(k)]
        # create the events outside the capture
        fork.record(s0)
        for e, st in zip(joins, self.streams):
            e.record(st)
        torch.cuda.synchronize(self.device)
        h = _lib.vp(s0.cuda_stream)
        _lib.call("me_graph_begin", h)
        try:
            fork.record(s0)
            for st in self.streams:
                st.wait_event(fork)
            for i, r in enumerate(self.recons):
                with torch.cuda.stream(self.streams[i % k]):
                    r.step()
            for e, st in zip(joins, self.streams):
                e.record(st)
                s0.wait_event(e)
            ex = C.c_void_p()
            _lib.call("me_graph_end", h, C.byref(ex))
        except Exception:
            _lib.call("me_graph_abort", h)
            raise
        self._exec = ex
        self._events = (fork, joins)
        self._capture_stream = s0

    def step(self):
        """One outer iteration of every volume: eager the first time, then by
        replaying the recorded graph on the current stream."""
        if self.steps_done == 0:
            self.step_eager()
            return
        if self._exec is None:
            self._capture()
        _lib.call("me_graph_launch", self._exec, _lib.vp(torch.cuda.current_stream(self.device).cuda_stream))
        self.steps_done += 1

    def run(self, n_iter):
        for _ in range(int(n_iter)):
            self.step()
        return [r.image() for r in self.recons], [r.t for r in self.recons]

    def close(self):
        if self._exec is not None:
            _lib.call("me_graph_destroy", self._exec)
            self._exec = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
