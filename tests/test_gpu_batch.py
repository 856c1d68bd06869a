"""Batched stepping (fp_batch_*): many independent scenes in one launch per
pass, a frame captured into a HIP graph and replayed.

Bars:
  * every scene of a batch is bit-identical, frame by frame, to the same
    scene stepped alone (same kernels, same per-scene arithmetic; nothing
    crosses scenes);
  * graph replay is bit-identical to direct launches, with host work and
    synchronisations between the replays;
  * membership changes between frames (remove, add) leave the remaining
    scenes bit-identical to stepping them alone.
"""
import numpy as np
import pytest

from tests._gpu import log

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
SUB = 5
KEYS = ("pos", "prev", "vel", "active")
TUNE_FRAMES = 4  # the first frames time the contact lane counts 4, 8, 16, 32 (launched directly)


@pytest.fixture(scope="module", autouse=True)
def _lib():
    from fluidphysics_amd import _abi
    _abi.load_library()


def _scenes(idx, frames):
    from fluidphysics_amd import scenes
    out = []
    for k in idx:
        sc = scenes.batched_scene(k, frames=frames)
        # bring the slice strokes into the short test horizon
        for j, w in enumerate(sc.wedges):
            w.frame = 3 + 5 * j + k % 4
        out.append(sc)
    return out


def _cuts(s, sc, frame):
    for w in sc.wedges:
        if w.frame == frame:
            s.cut(*w.o, *w.d1, *w.d2)
    g = getattr(sc, "test_grab", None)
    if g is not None:
        f0, f1, ids = g
        if frame == f0:
            cx, cy, cz = sc.cube[:3]
            s.grabs = [{"i": int(i), "x": cx + 0.3, "y": cy + 0.2 + 0.01 * k, "z": cz} for k, i in enumerate(ids)]
        elif frame == f1:
            s.grabs = []


def _snap(s):
    st = s.read_state()
    c = s.read_constraints()
    return {k: st[k].copy() for k in KEYS} | {"rest": c["rest"], "alive": c["alive"]}


def _same(a, b, label):
    for k in a:
        if not np.array_equal(a[k], b[k]):
            bad = np.flatnonzero(a[k].ravel() != b[k].ravel())
            raise AssertionError("%s: %s differs at %d entries (first %d)" % (label, k, bad.size, bad[0]))


def _alone(sc, frames):
    from fluidphysics_amd import from_scene
    s = from_scene(sc)
    snaps = []
    for f in range(frames):
        _cuts(s, sc, f)
        s.step(DT, SUB)
        snaps.append(_snap(s))
    s.close()
    return snaps


def test_batch_bit_identical_to_scenes_alone():
    from fluidphysics_amd import Batch, from_scene
    frames = 12
    scs = _scenes([0, 1, 2, 3, 5], frames)
    log("case batch vs alone: %s" % [sc.name for sc in scs])
    ref = [_alone(sc, frames) for sc in scs]
    b = Batch()
    ss = [from_scene(sc) for sc in scs]
    for s in ss:
        b.add(s)
    for f in range(frames):
        for s, sc in zip(ss, scs):
            _cuts(s, sc, f)
        b.step(DT, SUB)
        for k, s in enumerate(ss):
            _same(_snap(s), ref[k][f], "%s frame %d" % (scs[k].name, f))
    inf = b.info()
    assert inf.scenes == len(scs) and inf.frames == frames
    assert inf.graph_launches == frames - TUNE_FRAMES
    assert inf.contact_lanes in (4, 8, 16, 32)
    b.close()


def test_graph_replay_equals_direct_launches():
    from fluidphysics_amd import Batch, from_scene
    frames = 10
    scs = _scenes([6, 7, 8], frames)
    log("case graph vs direct: %s" % [sc.name for sc in scs])
    runs = []
    for use_graph in (False, True):
        b = Batch()
        b.graph(use_graph)
        ss = [from_scene(sc) for sc in scs]
        for s in ss:
            b.add(s)
        snaps = []
        for f in range(frames):
            for s, sc in zip(ss, scs):
                _cuts(s, sc, f)
            b.step(DT, SUB)
            # eager device work between replays: reads, a constraint
            # read-back, a fluid count
            snaps.append([_snap(s) for s in ss])
            ss[0].fluidCount()
        inf = b.info()
        assert inf.graph_launches == (frames - TUNE_FRAMES if use_graph else 0)
        runs.append(snaps)
        b.close()
    for f in range(frames):
        for k in range(len(scs)):
            _same(runs[1][f][k], runs[0][f][k], "%s frame %d" % (scs[k].name, f))


def test_membership_changes_and_mixed_parameters():
    """A scene removed mid-run and another added keep every member identical
    to stepping it alone; members differ in every parameter but
    constraint_iters (here 2, an even count of Jacobi sweeps), and one has a
    grab, which runs the frame without the graph."""
    from fluidphysics_amd import Batch, from_scene, scenes
    frames = 8
    scs = _scenes([9, 10, 11], frames)
    plain = scenes.plain_cpu_scene(frames=frames)
    plain.wedges[0].frame = 4
    for sc in scs:
        sc.params["constraint_iters"] = 2
    plain.params["constraint_iters"] = 2
    scs[0].test_grab = (2, 6, list(range(8)))
    log("case membership: %s + %s" % ([sc.name for sc in scs], plain.name))
    ref = {sc.name: _alone(sc, frames) for sc in scs}
    # the plain scene joins at frame 3: its alone run starts there too
    ref_plain = _alone(plain, frames - 3)
    b = Batch()
    ss = {sc.name: from_scene(sc) for sc in scs}
    for s in ss.values():
        b.add(s)
    gone = ss[scs[1].name]
    p = None
    for f in range(frames):
        if f == 3:
            p = from_scene(plain)
            b.add(p)
        if f == 5:
            b.remove(gone)
        members = [(sc, ss[sc.name]) for sc in scs if not (f >= 5 and ss[sc.name] is gone)]
        for sc, s in members:
            _cuts(s, sc, f)
        if p is not None:
            _cuts(p, plain, f - 3)
        b.step(DT, SUB)
        for sc, s in members:
            _same(_snap(s), ref[sc.name][f], "%s frame %d" % (sc.name, f))
        if p is not None:
            _same(_snap(p), ref_plain[f - 3], "%s frame %d" % (plain.name, f - 3))
    # the removed scene still steps alone, from where it left the batch
    _cuts(gone, scs[1], 5)
    gone.step(DT, SUB)
    _same(_snap(gone), ref[scs[1].name][5], "%s alone after removal" % scs[1].name)
    b.close()


def test_contact_lane_counts_give_identical_results():
    """The solid contact pass adds its per-cell partials by one fixed tree
    whatever its lane count: the tuning frames (each a different count) and
    the tuned frames give the same bits as a batch pinned to each count."""
    from fluidphysics_amd import Batch, from_scene
    frames = 6
    scs = _scenes([16, 17], frames)
    log("case contact lanes: %s" % [sc.name for sc in scs])
    runs = []
    for pin in (None, 4, 8, 16, 32):
        b = Batch()
        if pin:
            b.contact_lanes(pin)
        ss = [from_scene(sc) for sc in scs]
        for s in ss:
            b.add(s)
        snaps = []
        for f in range(frames):
            for s, sc in zip(ss, scs):
                _cuts(s, sc, f)
            b.step(DT, SUB)
            snaps.append([_snap(s) for s in ss])
        if pin:
            assert b.info().contact_lanes == pin
        runs.append(snaps)
        b.close()
    for r in runs[1:]:
        for f in range(frames):
            for k in range(len(scs)):
                _same(r[f][k], runs[0][f][k], "%s frame %d" % (scs[k].name, f))


def test_batch_rejects_mismatched_constraint_iters():
    from fluidphysics_amd import Batch, FpError, from_scene
    scs = _scenes([12, 13], 4)
    scs[1].params["constraint_iters"] = 4
    b = Batch()
    for sc in scs:
        b.add(from_scene(sc))
    with pytest.raises(FpError, match="constraint_iters"):
        b.step(DT, SUB)
    b.close()


def test_batch_profile_reports_every_pass():
    from fluidphysics_amd import Batch, from_scene
    scs = _scenes([14, 15], 4)
    b = Batch()
    for sc in scs:
        b.add(from_scene(sc))
    b.profile(True)
    for _ in range(2):
        b.step(DT, SUB)
    prof = b.profile_read()
    # bounds and velocity end each particle's substep inside the last
    # constraint iteration (solids) and the displacement (fluid)
    for name in ("integrate", "scan_apply", "scatter", "rank", "contacts", "resort", "density", "displace",
                 "constraints", "xsph"):
        assert name in prof and prof[name][0] >= 2 * SUB, (name, prof.get(name))
    assert b.info().graph_launches == 0
    b.close()


def test_forked_frames_equal_serial_frames():
    """A captured frame of a group of 131,072 particle slots or more runs the
    solid chain (solid contacts, constraint iterations) on a second stream
    beside the fluid chain; eager frames run them one after the other.  The
    two must give the same bits (the chains share no buffer they write)."""
    from fluidphysics_amd import Batch, from_scene, scenes
    frames = 9
    scs, slots = [], 0
    for k in range(40, 80):
        sc = _scenes([k], frames)[0]
        scs.append(sc)
        slots += sc.cube[3] ** 3 + sc.max_fluid
        if slots >= 160000:
            break
    assert slots >= 131072  # the side stream is used for this group
    log("case forked vs serial frames: %d scenes, %d slots" % (len(scs), slots))
    runs = []
    for use_graph in (True, False):
        b = Batch()
        b.graph(use_graph)
        ss = [from_scene(sc) for sc in scs]
        for s in ss:
            b.add(s)
        snaps = {}
        for f in range(frames):
            for s, sc in zip(ss, scs):
                _cuts(s, sc, f)
            b.step(DT, SUB)
            if f in (TUNE_FRAMES + 1, frames - 1):
                snaps[f] = [_snap(s) for s in ss]
        assert b.info().graph_launches == (frames - TUNE_FRAMES if use_graph else 0)
        runs.append(snaps)
        b.close()
    for f in runs[0]:
        for k in range(len(scs)):
            _same(runs[0][f][k], runs[1][f][k], "%s frame %d" % (scs[k].name, f))


@pytest.mark.parametrize("iters", [1, 2, 3, 4])
def test_riding_solid_chain_equals_its_own_launches(iters):
    """A group below 131,072 particle slots runs its solid chain inside the
    fluid chain's launches: the solid contacts beside the fluid contacts, the
    last three constraint iterations beside the fluid rank, the density and
    the displacement, earlier ones on launches of their own.  A larger group
    runs every solid pass on its own launch (one stream when eager, a second
    stream in a captured frame).  Scenes stepped alone (riding) equal the same
    scenes in such a batch, bit for bit, whatever the iteration count."""
    from fluidphysics_amd import Batch, from_scene
    frames = TUNE_FRAMES + 3
    scs, slots = [], 0
    for k in range(40, 80):
        sc = _scenes([k], frames)[0]
        sc.params["constraint_iters"] = iters
        scs.append(sc)
        slots += sc.cube[3] ** 3 + sc.max_fluid
        if slots >= 140000:
            break
    assert slots >= 131072
    log("case riding solid chain, %d iterations: %d scenes, %d slots" % (iters, len(scs), slots))
    picks = [0, len(scs) - 1]
    ref = {k: _alone(scs[k], frames) for k in picks}
    for use_graph in (True, False):
        b = Batch()
        b.graph(use_graph)
        ss = [from_scene(sc) for sc in scs]
        for s in ss:
            b.add(s)
        for f in range(frames):
            for s, sc in zip(ss, scs):
                _cuts(s, sc, f)
            b.step(DT, SUB)
            if f >= TUNE_FRAMES:
                for k in picks:
                    _same(_snap(ss[k]), ref[k][f], "%s frame %d (graph %s)" % (scs[k].name, f, use_graph))
        b.close()


def test_unconstrained_scene_finishes_its_solids():
    """A scene with solids but no constraints (a one-particle cube) has no
    constraint iteration to end its solids' substep: kg_bounds_velocity does
    (solver.js:389-427, :219-234).  Alone and in a batch beside a constrained
    scene it gives the same bits, and its solid comes to rest on the floor
    at its radius."""
    from fluidphysics_amd import Batch, Solver, from_scene
    frames = 60

    def lone():
        s = Solver(512)
        s.addSolidCube(0.3, 1.0, -0.2, 1, 0.18)
        return s

    log("case unconstrained scene alone and in a batch")
    a = lone()
    ref = []
    for _ in range(frames):
        a.step(DT, SUB)
        ref.append(_snap(a))
    r = a.solidRadius
    y = ref[-1]["pos"][1]
    assert abs(y - r) < 1e-6 and np.array_equal(ref[-1]["pos"], ref[-2]["pos"])
    a.close()

    other = _scenes([9], frames)[0]
    ref_other = _alone(other, frames)
    b = Batch()
    s1, s2 = lone(), from_scene(other)
    b.add(s1)
    b.add(s2)
    for f in range(frames):
        _cuts(s2, other, f)
        b.step(DT, SUB)
        _same(_snap(s1), ref[f], "one-particle cube frame %d" % f)
        if f % 10 == 9:
            _same(_snap(s2), ref_other[f], "%s frame %d" % (other.name, f))
    b.close()


def test_captured_frame_survives_membership_changes():
    """Members join and leave between replayed frames: the frame captured at
    the start keeps replaying (its kernels read the scene table and the work
    lists' counts on the device), and every member stays bit-identical to
    stepping it alone."""
    from fluidphysics_amd import Batch, from_scene
    frames = 14
    scs = _scenes([40, 41, 42, 43, 44], frames)
    log("case graph across membership changes: %s" % [sc.name for sc in scs])
    joins = {scs[3].name: 5, scs[4].name: 9}
    leaves = {scs[1].name: 7}
    ref = {sc.name: _alone(sc, frames - joins.get(sc.name, 0)) for sc in scs}
    b = Batch()
    b.contact_lanes(8)  # no tuning frames
    ss = {}
    for sc in scs[:3]:
        ss[sc.name] = from_scene(sc)
        b.add(ss[sc.name])
    for f in range(frames):
        for sc in scs:
            if joins.get(sc.name) == f:
                ss[sc.name] = from_scene(sc)
                b.add(ss[sc.name])
            if leaves.get(sc.name) == f:
                b.remove(ss.pop(sc.name))
        live = [sc for sc in scs if sc.name in ss]
        for sc in live:
            _cuts(ss[sc.name], sc, f - joins.get(sc.name, 0))
        b.step(DT, SUB)
        for sc in live:
            _same(_snap(ss[sc.name]), ref[sc.name][f - joins.get(sc.name, 0)], "%s frame %d" % (sc.name, f))
    inf = b.info()
    log("  graph launches %d, captures %d over %d frames" % (inf.graph_launches, inf.graph_captures, frames))
    # the first frame runs eagerly; every later one replays, and the
    # membership changes (5 -> 3 -> 4 -> 3 -> 4 scenes of similar sizes) do
    # not force a new capture each
    assert inf.graph_launches == frames - 1
    assert inf.graph_captures <= 2
    b.close()


def test_launch_shape_tuning_keeps_results():
    """The batch's launch-shape tuner runs every candidate (the density
    kernel's LDS staging on and off, XCD runs of 16, 4 and 1 items) over
    captured frames: every frame stays bit-identical to the scenes stepped
    alone, the batch settles on one shape, and a second batch of the same
    signature takes it from the cache without timing frames."""
    from fluidphysics_amd import Batch, from_scene
    frames = 24
    scs = _scenes([4, 9, 10, 11], frames)
    log("case launch-shape tuning: %s" % [sc.name for sc in scs])
    ref = [_alone(sc, frames) for sc in scs]
    b = Batch()
    ss = [from_scene(sc) for sc in scs]
    for s in ss:
        b.add(s)
    seen = set()
    for f in range(frames):
        for s, sc in zip(ss, scs):
            _cuts(s, sc, f)
        b.step(DT, SUB)
        inf = b.info()
        seen.add((inf.density_stage, inf.xcd_run))
        for k, s in enumerate(ss):
            _same(_snap(s), ref[k][f], "%s frame %d" % (scs[k].name, f))
    inf = b.info()
    log("  shapes run %s; settled on staging %d, XCD run %d" % (sorted(seen), inf.density_stage, inf.xcd_run))
    assert inf.shape_tuned == 1
    assert len(seen) == 6
    choice = (inf.density_stage, inf.xcd_run)
    b.close()
    b2 = Batch()
    ss2 = [from_scene(sc) for sc in scs]
    for s in ss2:
        b2.add(s)
    for f in range(TUNE_FRAMES + 2):
        b2.step(DT, SUB)
    inf2 = b2.info()
    assert inf2.shape_tuned == 1 and (inf2.density_stage, inf2.xcd_run) == choice
    b2.close()
