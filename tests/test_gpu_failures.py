"""Failure detection and recovery, forced from the host (every kernel exits
normally; the faults are non-finite numbers in a scene's own state): the
per-scene non-finite flag, a restore clearing it, and the serving loop
retiring failed and cancelled requests while the others step on.

Bars:
  * a NaN fluid particle emitted into one scene of a batch raises that
    scene's flag only, and the other scenes are bit-identical, every frame,
    to the same scenes stepped in a batch without it;
  * a solid pulled to +inf by a grab raises its solver's flag; restoring a
    checkpoint taken before the fault clears the flag, and the scene then
    steps bit-identically to the frames first stepped from that checkpoint;
  * serving (batch and per-scene paths): a request whose scene turns
    non-finite is retired as failed after the frame that did it, a request
    cancelled during a frame is retired after that frame, a request
    cancelled before admission is never admitted, and every other request's
    frames are bit-identical to the same request served alone.
"""
import math

import numpy as np
import pytest

from tests._gpu import log

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
SUB = 5


@pytest.fixture(scope="module", autouse=True)
def _lib():
    from fluidphysics_amd import _abi
    _abi.load_library()


def _nan_particle(s):
    s.emit_batch(np.array([[math.nan, 0.5, 0.0]]), np.zeros((1, 3)), 0.01)


def test_nonfinite_flag_marks_only_its_scene():
    from fluidphysics_amd import Batch, from_scene, scenes
    scs = [scenes.batched_scene(k, frames=8) for k in (50, 51, 52)]
    log("case nonfinite in a batch: %s, NaN fluid into the middle one" % [sc.name for sc in scs])

    def run(inject, frames=4):
        ss = [from_scene(sc) for sc in scs]
        b = Batch()
        for s in ss:
            b.add(s)
        if inject:
            _nan_particle(ss[1])
        out, flags = [], []
        for _ in range(frames):
            b.step(DT, SUB)
            out.append([s.read_positions().copy() for s in ss])
            flags.append([s.nonfinite() for s in ss])
        b.close()
        for s in ss:
            s.close()
        return out, flags

    clean, clean_flags = run(False)
    hurt, hurt_flags = run(True)
    assert all(f == [False, False, False] for f in clean_flags), clean_flags
    assert all(f == [False, True, False] for f in hurt_flags), hurt_flags
    for k in range(len(clean)):
        for j in (0, 2):
            assert np.array_equal(hurt[k][j], clean[k][j]), (k, j)
        assert np.isnan(hurt[k][1]).any()
    log("  flag raised in scene 1 only; scenes 0 and 2 bit-identical for %d frames" % len(clean))


def test_grab_to_infinity_flags_and_restore_clears():
    from fluidphysics_amd import from_scene, scenes
    sc = scenes.batched_scene(60, frames=8)
    log("case nonfinite solid: %s, solid 0 grabbed towards +inf" % sc.name)
    s = from_scene(sc)
    s.step(DT, SUB)
    s.checkpoint(0)
    ref = []
    for _ in range(2):
        s.step(DT, SUB)
        ref.append(s.read_positions().copy())
    assert not s.nonfinite()
    s.restore(0)
    s.grabs = [{"i": 0, "x": math.inf, "y": 1.0, "z": 0.0}]
    s.step(DT, SUB)
    assert s.nonfinite()
    assert not np.isfinite(s.read_positions()[: s.solidCount]).all()
    assert s.info().nonfinite == 1
    s.grabs = []
    s.restore(0)
    assert not s.nonfinite() and s.info().nonfinite == 0
    for k in range(2):
        s.step(DT, SUB)
        assert np.array_equal(s.read_positions(), ref[k]), k
        assert not s.nonfinite()
    s.close()


@pytest.mark.parametrize("mode", ["batch", "per_scene"])
def test_serving_retires_failed_and_cancelled_requests(mode):
    from fluidphysics_amd.scenes import Request
    from fluidphysics_amd.serving import Server
    H = 6
    reqs = [Request(id=k, arrival=k % 3, index=120 + k, seed=5, horizon=H) for k in range(6)]
    BAD, CANCEL, DROP = 1, 2, 5
    log("case serving failures (%s): request %d turns NaN, %d cancelled at frame 2, %d cancelled before admission"
        % (mode, BAD, CANCEL, DROP))

    def collect(target):
        frames = {}

        def on_frame(req, k, pos):
            if req.id == target:
                frames[k] = pos.copy()
        return frames, on_frame

    alone = {}
    for r in reqs:
        if r.id in (BAD, CANCEL, DROP):
            continue
        frames, cb = collect(r.id)
        rep = Server(mode=mode, on_frame=cb).run([r], by_frame=True)
        assert rep.completed == 1 and rep.failed == 0
        alone[r.id] = frames

    mix = {r.id: {} for r in reqs}
    status = {}
    server = None

    def on_frame(req, k, pos):
        mix[req.id][k] = pos.copy()
        if req.id == CANCEL and k == 2:
            server.cancel(CANCEL)

    def on_admit(req, s):
        if req.id == BAD:
            _nan_particle(s)

    def on_retire(req, st, frames):
        status[req.id] = (st, frames)

    server = Server(mode=mode, on_frame=on_frame, on_admit=on_admit, on_retire=on_retire)
    server.cancel(DROP)
    rep = server.run(reqs, by_frame=True)
    log("  %s" % rep.summary())
    assert status[BAD] == ("failed", 1), status
    assert status[CANCEL] == ("cancelled", 3), status
    assert status[DROP] == ("cancelled", 0), status
    assert rep.failed == 1 and rep.cancelled == 2 and rep.completed == len(alone)
    assert sorted(mix[DROP]) == []
    for rid, frames in alone.items():
        assert status[rid] == ("done", H)
        for k in range(H):
            assert np.array_equal(mix[rid][k], frames[k]), (rid, k)
