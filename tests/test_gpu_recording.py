"""Checkpoints, state blobs and recordings (the reference GPU page's snapshot /
restore and input-log replay, src/gpu/sim.js:494-547, src/gpu/main.js:308-390).

Bar: stepping on from a restored state -- device checkpoint, a blob loaded
into a fresh solver, or a recording replayed -- gives the same bits as
stepping on from the saved one, with slices, grabs, tears and emits in the
continuation.
"""
import json

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


def _scene(k=60):
    from fluidphysics_amd import scenes
    sc = scenes.batched_scene(k, frames=20)
    for j, w in enumerate(sc.wedges):
        w.frame = 3 + 2 * j
    return sc


def _state(s):
    st = s.read_state()
    c = s.read_constraints()
    rho, rhon = s.read_density()
    return {"pos": st["pos"], "prev": st["prev"], "vel": st["vel"], "active": st["active"], "group": st["group"],
            "rest": c["rest"], "alive": c["alive"], "rho": rho, "fluid": np.array([s.fluidCount()])}


def _same(a, b, label):
    for k in a:
        assert np.array_equal(a[k], b[k]), "%s: %s differs" % (label, k)


def _continue(s, sc, frames, f0=0):
    """The continuation: slices, a grab of eight solids that tears, an emit."""
    for f in range(f0, f0 + frames):
        for w in sc.wedges:
            if w.frame == f:
                s.cut(*w.o, *w.d1, *w.d2)
        if f == f0 + 1:
            cx, cy, cz = sc.cube[:3]
            s.grabs = [{"i": i, "x": cx + 1.5, "y": cy + 1.0, "z": cz} for i in range(8)]
        if f == f0 + 4:
            s.grabs = []
        if f == f0 + 2:
            p = np.array([[0.2 + 0.05 * k, 1.5, 0.1] for k in range(16)])
            s.emit_batch(p, np.zeros_like(p), 12 * 0.03 ** 3)
        s.step(DT, SUB)


def test_checkpoint_restore_replays_bit_identically():
    from fluidphysics_amd import from_scene
    sc = _scene()
    log("case checkpoint: %s" % sc.name)
    s = from_scene(sc)
    for _ in range(4):
        s.step(DT, SUB)
    s.checkpoint(0)
    _continue(s, sc, 8)
    a = _state(s)
    assert int(a["alive"].sum()) < len(a["alive"])  # the continuation tore or cut
    s.restore(0)
    s.grabs = []
    _continue(s, sc, 8)
    _same(_state(s), a, "after restore")
    s.close()


def test_state_blob_into_a_fresh_solver():
    from fluidphysics_amd import FpError, from_scene, scenes
    sc = _scene()
    log("case state blob: %s" % sc.name)
    s = from_scene(sc)
    for _ in range(5):
        s.step(DT, SUB)
    blob = s.save_state()
    _continue(s, sc, 6, f0=5)
    a = _state(s)
    t = from_scene(sc)  # a fresh solver of the same shape, other state
    t.load_state(blob)
    _continue(t, sc, 6, f0=5)
    _same(_state(t), a, "blob")
    # a scene of another shape refuses the blob
    other = from_scene(scenes.batched_scene(61, frames=4))
    with pytest.raises(FpError, match="another shape"):
        other.load_state(blob)
    for x in (s, t, other):
        x.close()


def test_recording_replays_to_the_same_state():
    from fluidphysics_amd import from_scene
    from fluidphysics_amd.recording import Recorder, replay
    sc = _scene(62)
    log("case recording: %s" % sc.name)
    s = from_scene(sc)
    for _ in range(3):
        s.step(DT, SUB)
    rec = Recorder(s, dict(max_fluid=sc.max_fluid, cube=sc.cube, params=sc.params_with()))
    cx, cy, cz = sc.cube[:3]
    for f in range(8):
        for w in sc.wedges:
            if w.frame == f:
                rec.cut(w.o, w.d1, w.d2)
        if f == 1:
            rec.grabs([{"i": i, "x": cx + 1.2, "y": cy + 0.8, "z": cz} for i in range(8)])
        if f == 3:
            rec.grabs([])
            rec.set("viscosity", 0.2)
        if f == 4:
            p = np.array([[0.3, 1.2 + 0.05 * k, 0.0] for k in range(10)])
            rec.emit_batch(p, np.tile([0.0, -1.0, 0.0], (10, 1)), 12 * 0.03 ** 3)
        rec.step(DT, SUB)
    a = _state(s)
    bundle = json.loads(json.dumps(rec.bundle()))  # survives a JSON round trip
    assert bundle["frames"] == 8
    frames = []
    r = replay(bundle, on_frame=lambda f, x: frames.append(f))
    assert frames == list(range(1, 9))
    _same(_state(r), a, "replay")
    s.close()
    r.close()


def test_state_blob_with_bad_cursor_or_flags():
    """A blob whose ring cursor lies outside [0, maxFluid) or whose substep
    count is negative is refused before any copy (the cursor addresses the
    next emit's slot); flag bytes other than 0 / 1 enter the device as 1."""
    import struct
    from fluidphysics_amd import FpError, from_scene
    sc = _scene(63)
    log("case blob validation: %s" % sc.name)
    s = from_scene(sc)
    s.step(DT, SUB)
    blob = bytearray(s.save_state())
    n = s.capacity
    # header: magic[8], version u32, capacity, solids, maxFluid, constraints,
    # fluid_next (i32 each), substeps_done i64, flags u32, reserved u32
    assert struct.unpack_from("<iiiii", blob, 12)[:4] == (n, s.solidCount, s.maxFluid, s.info().num_constraints)
    for cur, sub in ((-1, 0), (s.maxFluid, 0), (1 << 30, 0), (0, -5)):
        bad = bytearray(blob)
        struct.pack_into("<i", bad, 28, cur)
        struct.pack_into("<q", bad, 32, sub)
        with pytest.raises(FpError, match="out of range"):
            s.load_state(bytes(bad))
    # active flags of 7 load as 1
    act_off = 48 + n * (16 * 3 + 8)
    want = s.read_state()["active"].copy()
    bad = bytearray(blob)
    for i in range(n):
        if bad[act_off + i]:
            bad[act_off + i] = 7
    s.load_state(bytes(bad))
    assert np.array_equal(s.read_state()["active"], want)
    assert set(np.unique(want)) <= {0, 1}
    s.close()
