"""The substep's own cell sort (the two counting sorts every pass of the hot
path gathers through) against a brute-force sort of the same positions.

The reference grid (solver.js:174-184) is checked bit-exact elsewhere
through fp_build_grid; the substep does not use it.  It sorts the particles
by a cell key of its own, twice per substep: the grid order (every active
particle by its position after integrate and the grab pull) and the fluid
order (every active fluid particle by its post-contact position).  Here,
after one substep:
  * the grid order's slots hold exactly the active particles, at their
    integrated positions (within 1 float32 ulp-scale of the FP64
    integrate, solver.js:194-202), with the inverse mass in w;
  * the fluid order's slots hold exactly the active fluid, at the
    post-contact positions of the FP64 Jacobi oracle (within 1e-4 * h);
  * both orders equal a brute-force stable sort of those positions by
    (key, particle index), and the cell starts equal its key offsets:
    bit-exact.
"""
import numpy as np
import pytest

from tests._gpu import log, tol_report

pytestmark = pytest.mark.gpu

SUB = 1.0 / 60.0 / 5


@pytest.fixture(scope="module", autouse=True)
def _lib():
    from fluidphysics_amd import _abi
    _abi.load_library()


def _keys(pos, c):
    """The sort key of each position: cell = floor(p * inv_cell) per axis as a
    32-bit integer, key = (cx * mx + cy * my + cz) mod M (32-bit wrap)."""
    cell = np.floor(pos[:, :3].astype(np.float64) * c["inv_cell"])
    assert (np.abs(cell) < 2 ** 30).all()
    u = cell.astype(np.int64).astype(np.uint32)
    with np.errstate(over="ignore"):
        k = u[:, 0] * np.uint32(c["mx"]) + u[:, 1] * np.uint32(c["my"]) + u[:, 2]
    return (k & np.uint32(c["M"] - 1)).astype(np.int64)


def _check_sorted(c, label):
    """ids / start equal the brute-force stable sort of the slots' positions."""
    ids, pos = c["ids"], c["pos"]
    keys = _keys(pos, c)
    order = np.lexsort((ids, keys))
    assert np.array_equal(ids[order], ids) and np.array_equal(keys[order], keys), label + ": not (key, index) order"
    exp_start = np.searchsorted(keys, np.arange(c["M"] + 1), side="left")
    assert np.array_equal(c["start"], exp_start), label + ": cell starts differ"
    occupied = int((np.diff(exp_start) > 0).sum())
    log("  %s: %d slots in %d of %d keys, max %d per key" % (label, c["slots"], occupied, c["M"],
                                                          int(np.diff(exp_start).max())))


def _run(solvers, frames, dt=1.0 / 60.0, substeps=5):
    for _ in range(frames):
        for s in solvers:
            s.step(dt, substeps)


def _pre(s):
    return s.read_state(), s.read_constraints()


def _verify(s, pre, params, cube, max_fluid):
    """The cells of the substep that s just took from state `pre`."""
    from oracle.oracle import OracleSolver
    st0, cons = pre
    gravity = (params["gravity_x"], params["gravity"], params["gravity_z"])
    ns = cube[3] ** 3
    g0 = s.read_cells(0)
    g1 = s.read_cells(1)
    act = st0["active"].astype(bool)
    # grid order: every active particle once, at its integrated position
    assert np.array_equal(np.sort(g0["ids"]), np.nonzero(act)[0])
    p0 = st0["pos"].reshape(-1, 3).astype(np.float64)
    v0 = st0["vel"].reshape(-1, 3).astype(np.float64) + np.asarray(gravity) * SUB
    exp = (p0 + v0 * SUB)[g0["ids"]]
    e = tol_report("grid-order position", g0["pos"][:, :3], exp, 4e-6 * max(1.0, np.abs(exp).max()))
    assert e <= 4e-6 * max(1.0, np.abs(exp).max())
    assert np.array_equal(g0["pos"][:, 3], st0["inv_mass"][g0["ids"]])
    _check_sorted(g0, "grid order")
    # fluid order: every active fluid particle once, at its post-contact position
    fl = act.copy()
    fl[:ns] = False
    assert np.array_equal(np.sort(g1["ids"]), np.nonzero(fl)[0])
    wbits = g1["pos"][:, 3].view(np.uint32)
    assert np.array_equal((wbits & np.uint32(0x3FFFFFFF)).astype(np.int64), g1["ids"])
    o = OracleSolver(params, max_fluid, "f64", "jacobi")
    o.add_solid_cube(*cube)
    o.write_state(st0["pos"], st0["prev"], st0["vel"], st0["inv_mass"], st0["active"])
    o.write_constraints(cons["rest"].astype(np.float64), cons["alive"])
    o.substep_until_contacts(SUB)
    oc = o.read_state64()["pos"].reshape(-1, 3)[g1["ids"]]
    bound = 1e-4 * params["fluid_h"]
    e = tol_report("fluid-order position vs FP64", g1["pos"][:, :3], oc, bound)
    assert e <= bound
    _check_sorted(g1, "fluid order")


@pytest.mark.parametrize("scene_name,frames", [("plain_cpu", 0), ("plain_cpu", 90), ("single_gpu", 20),
                                               ("large", 8)])
def test_cell_sort_equals_brute_force(scene_name, frames):
    from fluidphysics_amd import from_scene, scenes
    log("case cell sort %s frames=%d" % (scene_name, frames))
    sc = scenes.SCENES[scene_name]()
    s = from_scene(sc)
    with pytest.raises(Exception):
        s.read_cells(0)  # nothing sorted before the first substep
    _run([s], frames, sc.dt, sc.substeps)
    pre = _pre(s)
    s.step(SUB, 1)
    _verify(s, pre, sc.params_with(), sc.cube, sc.max_fluid)


def test_cell_sort_of_batched_members_equals_brute_force():
    """Members of a batch (one launch per pass over all scenes, tilted
    gravity) sort like scenes alone."""
    from fluidphysics_amd import Batch, from_scene, scenes
    log("case cell sort batched scenes")
    scs = [scenes.batched_scene(k) for k in (0, 5, 11)]
    b = Batch()
    members = [from_scene(sc) for sc in scs]
    for m in members:
        b.add(m)
    for _ in range(6):
        b.step(scs[0].dt, scs[0].substeps)
    pres = [_pre(m) for m in members]
    b.step(SUB, 1)
    for sc, m, pre in zip(scs, members, pres):
        _verify(m, pre, sc.params_with(), sc.cube, sc.max_fluid)
    b.close()
