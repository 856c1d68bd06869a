"""Serving loop: independent scene requests stepped with continuous batching.

A request (fluidphysics_amd.scenes.Request) asks for `horizon` frames of one
scene.  The loop admits waiting requests in arrival order while the device
bytes of the admitted scenes (fp_scene_bytes, the exact figure the solver
will hold) stay within the HBM budget, steps every admitted scene one frame
per iteration, returns each scene's positions after every frame (a snapshot
of the whole batch copied to pinned host memory once), and retires a request
when its horizon is reached, all at frame boundaries.  State stays resident
on the device between frames; memory of retired scenes returns to the
device pool (no hipFree) and is reused by the next admissions.

Two ways to run the same loop:
  mode="batch"      one Batch: one launch per pass over every scene, the
                    frame replayed from a HIP graph, one snapshot copy
  mode="per_scene"  every scene through its own Solver, one after another,
                    each read back on its own (the plain per-scene path)

Host work of a frame (admission, scene set-up, slices) overlaps the device
work of the previous one: the frame is queued, the host prepares the next,
then waits for the snapshot.

Failures and cancellation, at frame boundaries: a scene whose frame ended a
particle with a non-finite position or velocity (the solver's non-finite
flag, raised on the device and read from host memory) is retired as failed
after that frame is delivered; cancel(request_id) retires a running request
(or drops a waiting one) at the next boundary.  The other scenes step on
unchanged: a scene's arithmetic does not depend on the batch around it.

The reference steps one scene in the browser (src/main.js:396 calls
solver.step once per animation frame); this is its many-scene form.
"""
import collections
import math
import os
import time
from dataclasses import dataclass, field

from . import Batch, from_scene, scene_bytes, tune_cache_load, tune_cache_save


@dataclass
class Served:
    """Per-request record: admission time, frames delivered, completion
    (status "done", "failed" or "cancelled")."""
    request: object
    admitted: float = None
    frames: int = 0
    done: float = None
    bytes: int = 0
    status: str = None


@dataclass
class ServeReport:
    mode: str
    wall_s: float
    frames: int
    frame_ms: list = field(default_factory=list)   # per loop frame: dispatch -> snapshot on host
    queue_ms: list = field(default_factory=list)   # per admitted request: arrival -> admission
    particle_substeps: int = 0
    completed: int = 0
    failed: int = 0                                # retired after a non-finite frame
    cancelled: int = 0                             # retired or dropped by cancel()
    admitted: int = 0
    arrived: int = 0
    refused: int = 0
    peak_bytes: int = 0
    peak_scenes: int = 0
    budget_bytes: int = 0

    @staticmethod
    def _pct(xs, q):
        if not xs:
            return None
        xs = sorted(xs)
        return xs[min(len(xs) - 1, int(math.ceil(q * len(xs))) - 1)]

    def summary(self):
        return {
            "mode": self.mode,
            "wall_s": round(self.wall_s, 3),
            "frames": self.frames,
            "value": round(self.particle_substeps / self.wall_s, 1) if self.wall_s > 0 else 0.0,
            "unit": "particle-substeps/s",
            "frame_ms_p50": _r(self._pct(self.frame_ms, 0.50)),
            "frame_ms_p99": _r(self._pct(self.frame_ms, 0.99)),
            "queue_ms_p50": _r(self._pct(self.queue_ms, 0.50)),
            "queue_ms_p99": _r(self._pct(self.queue_ms, 0.99)),
            "arrived": self.arrived,
            "admitted": self.admitted,
            "completed": self.completed,
            "failed": self.failed,
            "cancelled": self.cancelled,
            "refused": self.refused,
            "peak_scenes": self.peak_scenes,
            "peak_bytes": self.peak_bytes,
            "budget_bytes": self.budget_bytes,
        }


def _r(x):
    return None if x is None else round(x, 3)


class Server:
    """Continuous-batching server of scene requests on one device.

    on_frame(request, frame_index, positions) is called for every delivered
    frame; `positions` is a (capacity, 3) float32 view that is reused two
    frames later (copy what must be kept).  on_admit(request, solver) is
    called once a request's solver is set up, before its first frame;
    on_retire(request, status, frames) when it leaves ("done", "failed",
    "cancelled")."""

    def __init__(self, budget_bytes=200e9, mode="batch", device=-1, dt=1.0 / 60.0, substeps=5,
                 on_frame=None, use_graph=True, on_admit=None, on_retire=None, tune_cache=None):
        if mode not in ("batch", "per_scene"):
            raise ValueError("mode must be 'batch' or 'per_scene'")
        self.budget = int(budget_bytes)
        self.mode = mode
        self.device = device
        self.dt = dt
        self.substeps = substeps
        self.on_frame = on_frame
        self.use_graph = use_graph
        self.on_admit = on_admit
        self.on_retire = on_retire
        self.tune_cache = tune_cache  # a tuning cache file: loaded before run() if present, saved after
        self._cancel = set()

    def cancel(self, request_id):
        """Retire request `request_id` at the next frame boundary (a waiting
        one is dropped before admission)."""
        self._cancel.add(request_id)

    def _admit(self, waiting, used, now, by_frame, report):
        """Head-of-line admission in arrival order within the byte budget.
        `waiting` holds (request, scene, bytes)."""
        out = []
        while waiting:
            req, sc, need = waiting[0]
            if req.id in self._cancel:
                waiting.popleft()
                self._cancel.discard(req.id)
                report.cancelled += 1
                if self.on_retire is not None:
                    self.on_retire(req, "cancelled", 0)
                continue
            if used + need > self.budget:
                if need > self.budget:
                    waiting.popleft()  # can never fit: refuse it rather than stall the queue
                    report.refused += 1
                    continue
                break
            waiting.popleft()
            s = from_scene(sc, device=self.device)
            if self.on_admit is not None:
                self.on_admit(req, s)
            rec = Served(request=req, admitted=now, bytes=need)
            report.queue_ms.append(0.0 if by_frame else max(0.0, now - req.arrival) * 1e3)
            out.append((rec, sc, s))
            used += need
        return out, used

    def run(self, requests, duration=None, by_frame=False, max_frames=None, clock=time.perf_counter):
        """Serve `requests` (sorted by arrival).  by_frame=False: arrivals are
        seconds on `clock` from the start and the loop stops at `duration`
        seconds; by_frame=True: arrivals are loop-frame indices (a
        deterministic mix for tests) and the loop runs until every request
        completed or max_frames."""
        if self.tune_cache and os.path.exists(self.tune_cache):
            tune_cache_load(self.tune_cache)
        report = ServeReport(mode=self.mode, wall_s=0.0, frames=0, budget_bytes=self.budget)
        pending = collections.deque(sorted(requests, key=lambda r: (r.arrival, r.id)))
        waiting = collections.deque()
        active = []          # (Served, scene, solver) in batch order
        used = 0
        batch = None
        if self.mode == "batch":
            batch = Batch(self.device)
            batch.graph(self.use_graph)
        frame = 0
        in_flight = None     # (dispatch time, members of the frame)
        t0 = clock()
        try:
            while True:
                now = clock() - t0
                stamp = frame if by_frame else now
                stop = (duration is not None and now >= duration) or (max_frames is not None and frame >= max_frames)
                if not stop:
                    while pending and pending[0].arrival <= stamp:
                        req = pending.popleft()
                        sc = req.scene()
                        waiting.append((req, sc, scene_bytes(sc.params_with(), sc.max_fluid, sc.cube[3])))
                        report.arrived += 1
                    admitted, used = self._admit(waiting, used, stamp, by_frame, report)
                else:
                    admitted = []
                if in_flight is not None:
                    t_disp, members = in_flight
                    self._deliver(batch, members, report)
                    report.frame_ms.append((clock() - t0 - t_disp) * 1e3)
                    report.frames += 1
                    report.particle_substeps += sum(s.capacity for _, _, s in members) * self.substeps
                    in_flight = None
                    keep = []
                    for m in active:
                        rec, sc, s = m
                        if rec.frames >= rec.request.horizon:
                            rec.status = "done"
                            report.completed += 1
                        elif s.nonfinite():
                            rec.status = "failed"
                            report.failed += 1
                        elif rec.request.id in self._cancel:
                            self._cancel.discard(rec.request.id)
                            rec.status = "cancelled"
                            report.cancelled += 1
                        else:
                            keep.append(m)
                            continue
                        rec.done = clock() - t0
                        if self.on_retire is not None:
                            self.on_retire(rec.request, rec.status, rec.frames)
                        if batch is not None:
                            batch.remove(s)
                        used -= rec.bytes
                        s.close()
                    active = keep
                if stop:
                    break
                for m in admitted:
                    if batch is not None:
                        batch.add(m[2])
                    active.append(m)
                    report.admitted += 1
                report.peak_scenes = max(report.peak_scenes, len(active))
                report.peak_bytes = max(report.peak_bytes, used)
                if not active:
                    if not pending and not waiting:
                        break
                    if not by_frame and pending:
                        time.sleep(max(0.0, min(0.05, pending[0].arrival - (clock() - t0))))
                    if by_frame:
                        frame += 1
                    continue
                for rec, sc, s in active:
                    for w in sc.wedges:
                        if w.frame == rec.frames:
                            s.cut(*w.o, *w.d1, *w.d2)
                t_disp = clock() - t0
                if batch is not None:
                    batch.step(self.dt, self.substeps, wait=False)
                else:
                    for _, _, s in active:
                        s.step(self.dt, self.substeps, wait=False)
                in_flight = (t_disp, list(active))
                frame += 1
        finally:
            report.wall_s = clock() - t0
            for _, _, s in active:
                if batch is not None:
                    batch.remove(s)
                s.close()
            if batch is not None:
                batch.close()
            if self.tune_cache:
                tune_cache_save(self.tune_cache)
        return report

    def _deliver(self, batch, members, report):
        if batch is not None:
            views = batch.positions()
            batch.sync()
        else:
            views = [s.read_positions() for _, _, s in members]
        for (rec, sc, s), pos in zip(members, views):
            if self.on_frame is not None:
                self.on_frame(rec.request, rec.frames, pos)
            rec.frames += 1
