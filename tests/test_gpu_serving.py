"""Serving: batch position snapshots, the admission estimate, the memory
pool, and the serving loop's results.

Bars:
  * a batch snapshot equals each scene's own read-back, bit for bit;
  * fp_scene_bytes equals the device bytes the solver then holds;
  * a request served alone and the same request served in a changing mix
    (arrivals and departures during its horizon, in the batch and in the
    per-scene path) deliver bit-identical positions every frame;
  * admission keeps the admitted scenes' bytes within the budget, in
    arrival order, and refuses a request larger than the budget.
"""
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


def test_snapshot_equals_read_back():
    from fluidphysics_amd import Batch, from_scene, scenes
    scs = [scenes.batched_scene(k, frames=8) for k in (20, 21, 22)]
    log("case snapshot: %s" % [s.name for s in scs])
    b = Batch()
    ss = [from_scene(sc) for sc in scs]
    for s in ss:
        b.add(s)
    first = None
    for f in range(3):
        b.step(DT, SUB)
        views = b.positions()
        assert len(views) == len(ss)
        for s, v in zip(ss, views):
            assert v.shape == (s.capacity, 3)
            assert np.array_equal(v, s.read_positions())
        if f == 0:
            first = [v.copy() for v in views]
            first_views = views
        if f == 1:
            # the previous snapshot stays readable until the second-next call
            for a, v in zip(first, first_views):
                assert np.array_equal(a, v)
    b.close()


def test_scene_bytes_is_what_the_solver_holds():
    from fluidphysics_amd import from_scene, pool_info, scene_bytes, scenes
    for k in (30, 31, 32, 33):
        sc = scenes.batched_scene(k, frames=4)
        est = scene_bytes(sc.params_with(), sc.max_fluid, sc.cube[3])
        s = from_scene(sc)
        assert s.info().device_bytes == est, (sc.name, s.info().device_bytes, est)
        s.close()
    # a destroyed scene's blocks are reused: the same scene again reserves nothing new
    sc = scenes.batched_scene(34, frames=4)
    s = from_scene(sc)
    s.close()
    r0, c0 = pool_info()
    s = from_scene(sc)
    r1, c1 = pool_info()
    assert r1 == r0 and c1 < c0
    s.close()


def _collect(target):
    frames = {}

    def on_frame(req, k, pos):
        if req.id == target:
            frames[k] = pos.copy()
    return frames, on_frame


def test_served_alone_equals_served_in_a_mix():
    from fluidphysics_amd.scenes import Request
    from fluidphysics_amd.serving import Server
    R = Request(id=1000, arrival=0, index=41, seed=5, horizon=14)
    sc = R.scene()
    for j, w in enumerate(sc.wedges):  # slices inside the short horizon
        w.frame = 4 + 3 * j
    R.scene = lambda: sc
    log("case serving alone vs mix: %s" % sc.name)
    alone, cb = _collect(R.id)
    rep = Server(mode="batch", on_frame=cb).run([R], by_frame=True)
    assert rep.completed == 1 and sorted(alone) == list(range(R.horizon))
    others = [Request(id=k, arrival=k % 7, index=100 + k, seed=5, horizon=2 + (k * 5) % 11) for k in range(9)]
    R2 = Request(id=1000, arrival=3, index=41, seed=5, horizon=14)
    R2.scene = lambda: sc
    for mode in ("batch", "per_scene"):
        mix, cb = _collect(R.id)
        rep = Server(mode=mode, on_frame=cb).run(others + [R2], by_frame=True)
        assert rep.completed == len(others) + 1, rep.summary()
        assert rep.peak_scenes >= 4
        for k in range(R.horizon):
            assert np.array_equal(mix[k], alone[k]), (mode, k)


def test_admission_within_budget_in_arrival_order():
    from fluidphysics_amd import scene_bytes
    from fluidphysics_amd.scenes import Request
    from fluidphysics_amd.serving import Server
    reqs = [Request(id=k, arrival=0, index=200 + k, seed=5, horizon=3 + k % 4) for k in range(10)]
    need = []
    for r in reqs:
        sc = r.scene()
        need.append(scene_bytes(sc.params_with(), sc.max_fluid, sc.cube[3]))
    budget = sorted(need)[-1] + sorted(need)[-2]   # two of the largest fit at once
    big = Request(id=99, arrival=0, index=300, seed=5, horizon=2)
    order = []

    def on_frame(req, k, pos):
        if k == 0:
            order.append(req.id)
    rep = Server(budget_bytes=budget, on_frame=on_frame).run(reqs, by_frame=True)
    log("admission: %s" % rep.summary())
    assert rep.completed == len(reqs)
    assert 0 < rep.peak_bytes <= budget
    assert order == [r.id for r in reqs]
    bsc = big.scene()
    big_need = scene_bytes(bsc.params_with(), bsc.max_fluid, bsc.cube[3])
    rep = Server(budget_bytes=big_need - 1).run([big], by_frame=True)
    assert rep.refused == 1 and rep.completed == 0 and rep.frames == 0
