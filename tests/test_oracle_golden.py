"""Pins the oracle (oracle/fp_oracle.c, float32 stores, Gauss-Seidel mode, one
sweep) bit for bit against the golden vectors the reference solver
src/solver.js wrote under node (tests/golden/make_golden.mjs).

What this pins that nothing else does: the Gauss-Seidel ordering of contacts
(src/solver.js:251-278) and constraints (:359-380), the silent pair cap
(:311), the exact double-compute / float32-store rounding sequence of a whole
step, tearing + union-find groups, and 300-frame trajectories.
"""
import numpy as np
import pytest

from tests import _golden as G


def _check_state(s, ref, keys=("pos", "prev", "vel", "active")):
    st = s.read_state()
    for k in keys:
        a, b = st[k], ref[k]
        if not np.array_equal(a.view(np.uint8), b.view(np.uint8)):
            bad = np.nonzero(a != b)[0]
            raise AssertionError("%s differs at %d entries, first %s: %r vs %r"
                                 % (k, bad.size, bad[:5], a[bad[:5]], b[bad[:5]]))


@pytest.mark.parametrize("name", ["plain_cpu", "ring_cap"])
def test_grid_initial_bit_exact(name):
    """_buildGrid (src/solver.js:174-184) on the initial state."""
    d, t = G.scenario(name), G.trace(name)
    s = G.make_oracle(d)
    inf = s.info()
    assert inf["table_size"] == t["tableSize"] and inf["capacity"] == t["capacity"]
    s.build_grid()
    cs, ce = s.read_grid()
    rcs, rce = G.load_grid0(name, t["tableSize"], t["capacity"])
    assert np.array_equal(cs, rcs)
    n = cs[-1]
    assert np.array_equal(ce[:n], rce[:n])


@pytest.mark.parametrize("name", ["plain_cpu", "ring_cap"])
def test_initial_state_and_one_substep(name):
    d, t = G.scenario(name), G.trace(name)
    s = G.make_oracle(d)
    st = s.read_state()
    assert G.sha(st["pos"]) == t["init"]["digests"]["pos"]
    assert G.sha(st["vel"]) == t["init"]["digests"]["vel"]
    s.step(1.0 / 60.0 / d["substeps"], 1)
    ref = G.load_state(name, "sub1", t["capacity"], t["solidCount"], t["numCons"])
    _check_state(s, ref)
    rho, rn = s.read_density()
    assert np.array_equal(rho, ref["rho"]) and np.array_equal(rn, ref["rho_near"])
    assert s.info()["pair_count"] == t["sub1"]["stats"]["pairs"]


def _replay(name, frames=None):
    d, t = G.scenario(name), G.trace(name)
    s = G.make_oracle(d)
    drv = G.Driver(d)
    nf = len(t["frames"]) if frames is None else frames
    for f in range(nf):
        drv.before_frame(s, f, lambda: s.read_state()["pos"], lambda: s.read_state()["group"])
        s.step(1.0 / 60.0, d["substeps"])
        st = s.read_state()
        ref = t["frames"][f]
        c = s.read_constraints()
        got = dict(pos=G.sha(st["pos"]), vel=G.sha(st["vel"]), active=G.sha(st["active"]),
                   alive=G.sha(c["alive"]), rest=G.sha(c["rest"]), group=G.sha(st["group"]))
        for k, v in got.items():
            assert v == ref["digests"][k], "%s frame %d: %s differs" % (name, f + 1, k)
        assert s.info()["pair_count"] == ref["stats"]["pairs"]
        assert int(c["alive"].sum()) == ref["stats"]["alive"]
    return s, drv, t


def test_plain_cpu_300_frames_bit_exact():
    """Every frame of the 300-frame plain scene, including the slice at
    frame 100, matches the reference bit for bit."""
    s, drv, t = _replay("plain_cpu")
    assert [n for _, n in drv.cuts] == [c["n"] for c in t["cuts"]]
    ref = G.load_state("plain_cpu", "f150", t["capacity"], t["solidCount"], t["numCons"])
    assert ref["alive"].sum() == t["frames"][149]["stats"]["alive"]


def test_grab_tear_bit_exact():
    """pickSolid, grab pull, tearing and the union-find groups."""
    s, drv, t = _replay("grab_tear")
    assert len(drv.picks) == 1
    (f, hit), = drv.picks
    assert hit[0] == t["picks"][0]["hit"]["id"] and hit[1] == t["picks"][0]["hit"]["t"]
    assert int(s.read_constraints()["alive"].sum()) < t["numCons"]  # the drag tore constraints


def test_ring_wrap_cull_pair_cap_bit_exact():
    """Ring overwrite, culling, the pair cap and the 2^53 hash path."""
    s, drv, t = _replay("ring_cap")
    assert t["sub1"]["stats"]["pairs"] == 32 * 256  # the cap bound in the first substep
    ref = G.load_state("ring_cap", "f30", t["capacity"], t["solidCount"], t["numCons"])
    _check_state(s, ref)
