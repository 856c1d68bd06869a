"""GPU parity: the HIP solver (through the C ABI) against the oracle and the
golden vectors of the reference solver.

Bars (from the accuracy contract of the path):
  * grid cellStart / cellEntries, post-cut alive flags, pick results:
    bit-exact;
  * positions after one substep from an identical state: within 1e-4 * h of
    the FP64 replay (h = the fluid interaction radius fluidH); velocities
    compared as v * dt_sub against the same bound (v itself is a difference
    quotient whose float32 resolution exceeds 1e-4*h/dt_sub);
  * 300 frames of the plain scene: time-mean kinetic energy, time-mean fluid
    density and final alive-constraint count within 1% of the reference
    (src/solver.js) trace.
"""
import numpy as np
import pytest

from tests import _golden as G
from tests._gpu import GpuAdapter, log, make_gpu, tol_report

pytestmark = pytest.mark.gpu

SUB = 1.0 / 60.0 / 5


def _load_lib_or_fail():
    from fluidphysics_amd import _abi
    _abi.load_library()


@pytest.fixture(scope="module", autouse=True)
def _lib():
    _load_lib_or_fail()


# ---------------------------------------------------------------- grid ----

@pytest.mark.parametrize("name", ["plain_cpu", "ring_cap"])
def test_grid_matches_reference_golden(name):
    log("case grid golden %s" % name)
    d, t = G.scenario(name), G.trace(name)
    s = make_gpu(d)
    assert s.info().table_size == t["tableSize"]
    s.build_grid()
    cs, ce = s.read_grid()
    rcs, rce = G.load_grid0(name, t["tableSize"], t["capacity"])
    assert np.array_equal(cs, rcs)
    assert np.array_equal(ce[: cs[-1]], rce[: cs[-1]])


def _scene_pair(scene_name, frames):
    """The scene's GPU solver after `frames` frames of its workload: the
    hose bursts (config 3) and the wedges (config 2) of those frames
    included."""
    from fluidphysics_amd import from_scene, scenes
    sc = scenes.SCENES[scene_name]()
    g = from_scene(sc)
    sp = 1.91 * sc.params_with()["fluid_radius"]
    for f in range(frames):
        if sc.hose:
            p, v = scenes.hose_burst(sc.hose, f, sp)
            g.emit_batch(p, v, sc.fluid_mass)
        for w in sc.wedges:
            if w.frame == f:
                g.cut(*w.o, *w.d1, *w.d2)
        g.step(sc.dt, sc.substeps)
    return sc, g


@pytest.mark.parametrize("scene_name,frames", [("single_gpu", 0), ("single_gpu", 20), ("large", 0), ("large", 10)])
def test_grid_matches_oracle_at_scale(scene_name, frames):
    """Bit-exact grid at the GPU-scale workloads, initial and mid-run."""
    from oracle.oracle import OracleSolver
    log("case grid oracle %s frames=%d" % (scene_name, frames))
    sc, g = _scene_pair(scene_name, frames)
    st = g.read_state()
    o = OracleSolver(sc.params_with(), sc.max_fluid, "f32", "jacobi")
    o.add_solid_cube(*sc.cube)
    o.write_state(st["pos"], st["prev"], st["vel"], st["inv_mass"], st["active"])
    o.build_grid()
    g.build_grid()
    cs, ce = g.read_grid()
    ocs, oce = o.read_grid()
    assert np.array_equal(cs, ocs)
    assert np.array_equal(ce[: cs[-1]], oce[: cs[-1]])


# ------------------------------------------------------- one substep ------

AMBIGUOUS_D = 2e-6  # |d - H| below this: float32 position rounding can flip the pair


def _ambiguous(o_state, H, params, max_fluid, cube, substep, grabs=None):
    """Fluid particles with a neighbour at |d - H| < AMBIGUOUS_D after the
    contact pass of the FP64 replay.  XSPH counts every pair inside H
    (solver.js:440-445), so whether such a pair is in the list is decided by
    the last bits of float32 positions -- for the reference's own
    Float32Array arithmetic as much as for the GPU -- and the velocity of
    both particles moves by ~|v_i - v_j|*0.3/m when it flips."""
    from scipy.spatial import cKDTree
    from oracle.oracle import OracleSolver
    o = OracleSolver(params, max_fluid, "f64", "jacobi")
    o.add_solid_cube(*cube)
    o.write_state(*o_state["state"])
    o.write_constraints(*o_state["cons"])
    if grabs is not None:
        o.set_grabs(*grabs)
    o.substep_until_contacts(substep)
    st = o.read_state()
    p = o.read_state64()["pos"].reshape(-1, 3)
    act = st["active"].astype(bool)
    act[: o.info()["solid_count"]] = False
    idx = np.nonzero(act)[0]
    tree = cKDTree(p[idx])
    pairs = tree.query_pairs(H + AMBIGUOUS_D, output_type="ndarray")
    d = np.linalg.norm(p[idx[pairs[:, 0]]] - p[idx[pairs[:, 1]]], axis=1)
    amb = pairs[np.abs(d - H) < AMBIGUOUS_D]
    out = np.zeros(len(act), bool)
    out[idx[amb.ravel()]] = True
    return out


def _compare_substep(g, o, H, label, amb=None, o32=None):
    st = g.read_state()
    ref = o.read_state64()
    ost = o.read_state()
    bound = 1e-4 * H
    assert np.array_equal(st["active"], ost["active"]), "active flags differ"
    act = st["active"].astype(bool)
    a3 = np.repeat(act, 3)
    ep = tol_report(label + " pos", st["pos"][a3], ref["pos"][a3], bound)
    keep = act if amb is None else act & ~amb
    k3 = np.repeat(keep, 3)
    if amb is not None:
        log("  %d of %d active particles have a pair within %.0e of d = H (excluded from the velocity bound)"
            % (int((amb & act).sum()), int(act.sum()), AMBIGUOUS_D))
    ev = tol_report(label + " vel*dt_sub", st["vel"][k3] * SUB, ref["vel"][k3] * SUB, bound)
    tol_report(label + " vel (literal)", st["vel"][k3], ref["vel"][k3], bound)
    tol_report(label + " vel*dt_sub, all", st["vel"][a3] * SUB, ref["vel"][a3] * SUB, bound)
    if o32 is not None:
        r32 = o32.read_state64()
        tol_report(label + " vel*dt_sub vs f32 oracle", st["vel"][a3] * SUB, r32["vel"][a3] * SUB, bound)
        tol_report(label + " pos vs f32 oracle", st["pos"][a3], r32["pos"][a3], bound)
    assert ep <= bound, "positions exceed 1e-4*h"
    assert ev <= bound, "velocities (v*dt_sub) exceed 1e-4*h"
    return ep, ev


def test_one_substep_strict_vs_fp64_reference_replay():
    """Initial plain scene (lattice at rest, fluid clear of the cube): the
    GS order of the reference and the Jacobi order coincide, so the GPU is
    held to the FP64 replay of the reference itself (Gauss-Seidel mode)."""
    log("case one-substep strict plain_cpu")
    d = G.scenario("plain_cpu")
    g = make_gpu(d, iters=3, omega=8.0)
    o = G.make_oracle(d, storage="f64", mode="gs", iters=1)
    g.step(SUB, 1)
    o.step(SUB, 1)
    _compare_substep(g, o, 0.19, "strict")


def _state_from_dump(name, suffix):
    t = G.trace(name)
    return G.load_state(name, suffix, t["capacity"], t["solidCount"], t["numCons"])


def test_one_substep_mid_run_vs_fp64_jacobi():
    """Frame-150 state of the reference run (cube resting in the fluid,
    sliced): one substep on the GPU vs the FP64 oracle in the Jacobi
    formulation; the gap to the reference's Gauss-Seidel order is reported."""
    log("case one-substep mid-run plain_cpu f150")
    d = G.scenario("plain_cpu")
    ref = _state_from_dump("plain_cpu", "f150")
    g = make_gpu(d, iters=3, omega=8.0)
    o = G.make_oracle(d, storage="f64", mode="jacobi", iters=3, jacobi_omega=8.0)
    og = G.make_oracle(d, storage="f64", mode="gs", iters=1)
    inv_mass = g.read_state()["inv_mass"]
    for s in (g, o, og):
        s.write_state(ref["pos"], ref["prev"], ref["vel"], inv_mass, ref["active"])
    g.write_constraints(ref["rest"].astype(np.float32), ref["alive"])
    o.write_constraints(ref["rest"], ref["alive"])
    og.write_constraints(ref["rest"], ref["alive"])
    g.step(SUB, 1)
    o.step(SUB, 1)
    og.step(SUB, 1)
    _compare_substep(g, o, 0.19, "vs jacobi")
    st = g.read_state()
    gs = og.read_state64()
    act = np.repeat(st["active"].astype(bool), 3)
    tol_report("gap to reference GS order", st["pos"][act], gs["pos"][act], 1e-4 * 0.19)


@pytest.mark.parametrize("iters", [1, 2])
def test_one_substep_few_iterations_vs_fp64_jacobi(iters):
    """The frame-150 state with one and two constraint iterations.  A single
    iteration reads its snapshot from the solid contacts' copy and finishes
    the solids into pos in the same launch; two iterations take the
    ping-pong buffers.  Both within 1e-4 h of the FP64 Jacobi oracle."""
    log("case one-substep f150, %d constraint iterations" % iters)
    d = G.scenario("plain_cpu")
    ref = _state_from_dump("plain_cpu", "f150")
    g = make_gpu(d, iters=iters, omega=8.0)
    o = G.make_oracle(d, storage="f64", mode="jacobi", iters=iters, jacobi_omega=8.0)
    inv_mass = g.read_state()["inv_mass"]
    for s in (g, o):
        s.write_state(ref["pos"], ref["prev"], ref["vel"], inv_mass, ref["active"])
    g.write_constraints(ref["rest"].astype(np.float32), ref["alive"])
    o.write_constraints(ref["rest"], ref["alive"])
    g.step(SUB, 1)
    o.step(SUB, 1)
    _compare_substep(g, o, 0.19, "%d iterations vs jacobi" % iters)


@pytest.mark.parametrize("scene_name,frames", [("single_gpu", 0), ("single_gpu", 30), ("large", 0), ("large", 12)])
def test_one_substep_at_scale_vs_fp64_jacobi(scene_name, frames):
    from oracle.oracle import OracleSolver
    log("case one-substep %s frames=%d" % (scene_name, frames))
    sc, g = _scene_pair(scene_name, frames)
    st = g.read_state()
    c = g.read_constraints()
    params = sc.params_with()
    state = dict(state=(st["pos"], st["prev"], st["vel"], st["inv_mass"], st["active"]),
                 cons=(c["rest"].astype(np.float64), c["alive"]))
    oracles = {}
    for storage in ("f64", "f32"):
        o = OracleSolver(params, sc.max_fluid, storage, "jacobi")
        o.add_solid_cube(*sc.cube)
        o.write_state(*state["state"])
        o.write_constraints(*state["cons"])
        o.step(SUB, 1)
        oracles[storage] = o
    amb = _ambiguous(state, params["fluid_h"], params, sc.max_fluid, sc.cube, SUB)
    g.step(SUB, 1)
    _compare_substep(g, oracles["f64"], params["fluid_h"], scene_name, amb=amb, o32=oracles["f32"])


# ------------------------------------------------------------- knife ------

def test_cut_bit_exact_plain():
    log("case cut plain_cpu f150")
    d = G.scenario("plain_cpu")
    ref = _state_from_dump("plain_cpu", "f150")
    g = make_gpu(d)
    o = G.make_oracle(d)
    inv_mass = g.read_state()["inv_mass"]
    for s in (g, o):
        s.write_state(ref["pos"], ref["prev"], ref["vel"], inv_mass, ref["active"])
    alive0 = np.ones_like(ref["alive"])
    g.write_constraints(None, alive0)
    o.write_constraints(None, alive0)
    from fluidphysics_amd import scenes
    rng = scenes.uniform01(11, 64 * 6).reshape(64, 6)
    cam = np.array([5.2, 3.6, 6.4])
    total = 0
    for k in range(64):
        a = np.array([rng[k, 0] - 0.5, 0.2 + 1.2 * rng[k, 1], rng[k, 2] - 0.5])
        b = np.array([rng[k, 3] - 0.5, 0.2 + 1.2 * rng[k, 4], rng[k, 5] - 0.5])
        d1 = (a - cam) / np.linalg.norm(a - cam)
        d2 = (b - cam) / np.linalg.norm(b - cam)
        ng = g.cut(*cam, *d1, *d2)
        no = o.cut(cam, d1, d2)
        assert ng == no, "cut %d: %d vs %d" % (k, ng, no)
        total += ng
    assert np.array_equal(g.read_constraints()["alive"], o.read_constraints()["alive"])
    assert total > 0
    # a degenerate wedge (d1 parallel to d2) cuts nothing
    assert g.cut(*cam, 0, 0, 1, 0, 0, 2) == 0


def test_pick_matches_oracle():
    log("case pick plain_cpu f150")
    d = G.scenario("plain_cpu")
    ref = _state_from_dump("plain_cpu", "f150")
    g = make_gpu(d)
    o = G.make_oracle(d)
    inv_mass = g.read_state()["inv_mass"]
    for s in (g, o):
        s.write_state(ref["pos"], ref["prev"], ref["vel"], inv_mass, ref["active"])
    cam = np.array([5.2, 3.6, 6.4])
    from fluidphysics_amd import scenes
    rng = scenes.uniform01(12, 40 * 3).reshape(40, 3)
    hits = 0
    for k in range(40):
        tgt = np.array([rng[k, 0] - 0.5, 0.3 + 0.8 * rng[k, 1], rng[k, 2] - 0.5])
        dd = (tgt - cam) / np.linalg.norm(tgt - cam)
        hg = g.pickSolid(*cam, *dd, 0.2)
        ho = o.pick_solid(cam, dd, 0.2)
        if ho is None:
            assert hg is None
        else:
            hits += 1
            assert hg is not None and hg["id"] == ho[0] and hg["t"] == ho[1]
    assert hits > 0


# ------------------------------------------------------ trajectories ------

def test_plain_scene_300_frames_statistics_within_1pct():
    """The solver as the drop-in creates it (library defaults: 3 Jacobi
    iterations, relaxation 8; only gravity set, as the scene does) against
    the reference's own 300-frame trace."""
    log("case 300 frames plain_cpu (library defaults)")
    d, t = G.scenario("plain_cpu"), G.trace("plain_cpu")
    g = make_gpu(d, iters=None, omega=None)
    assert g.params()["constraint_iters"] == 3 and g.params()["jacobi_omega"] == 8.0
    drv = G.Driver(d)
    ad = GpuAdapter(g)
    ke, rho = [], []
    for f in range(300):
        drv.before_frame(ad, f, lambda: g.read_state()["pos"], lambda: g.read_state()["group"])
        g.step(1.0 / 60.0, 5)
        st = g.read_state()
        r, _ = g.read_density()
        fs = G.frame_stats(st, t["solidCount"], r)
        ke.append(fs["ke"])
        rho.append(fs["meanRho"])
        assert np.isfinite(st["pos"]).all()
    rke = np.mean([f["stats"]["ke"] for f in t["frames"]])
    rrho = np.mean([f["stats"]["meanRho"] for f in t["frames"]])
    ral = t["frames"][-1]["stats"]["alive"]
    alive = int(g.read_constraints()["alive"].sum())
    log("  mean KE %.4f vs %.4f (%+.3f%%)" % (np.mean(ke), rke, 100 * (np.mean(ke) / rke - 1)))
    log("  mean rho %.5f vs %.5f (%+.3f%%)" % (np.mean(rho), rrho, 100 * (np.mean(rho) / rrho - 1)))
    log("  alive %d vs %d; cut %s vs %s" % (alive, ral, drv.cuts, [c["n"] for c in t["cuts"]]))
    assert abs(np.mean(ke) / rke - 1) < 0.01
    assert abs(np.mean(rho) / rrho - 1) < 0.01
    assert abs(alive / ral - 1) < 0.01


def test_grab_tear_and_groups():
    """The reference's grab scenario: tearing happens and the component labels
    are the union-find roots of the alive constraints (src/solver.js:488-499)."""
    log("case grab_tear")
    d = G.scenario("grab_tear")
    g = make_gpu(d, iters=3, omega=8.0)
    drv = G.Driver(d)
    ad = GpuAdapter(g)
    for f in range(len(G.trace("grab_tear")["frames"])):
        drv.before_frame(ad, f, lambda: g.read_state()["pos"], lambda: g.read_state()["group"])
        g.step(1.0 / 60.0, 5)
    c = g.read_constraints()
    assert c["alive"].sum() < len(c["alive"])
    grp = g.read_state()["group"]
    o = G.make_oracle(d)
    o.write_constraints(None, c["alive"])
    o.recompute_groups()
    assert np.array_equal(grp, o.read_state()["group"])
    assert drv.picks and drv.picks[0][1] is not None


def test_ring_wrap_and_initial_state_bit_exact():
    """emitFluid ring overwrite (src/solver.js:140-141) on the device matches
    the reference's initial state bit for bit."""
    log("case ring_cap initial")
    d, t = G.scenario("ring_cap"), G.trace("ring_cap")
    g = make_gpu(d)
    st = g.read_state()
    assert G.sha(st["pos"]) == t["init"]["digests"]["pos"]
    assert G.sha(st["vel"]) == t["init"]["digests"]["vel"]
    assert G.sha(st["active"]) == t["init"]["digests"]["active"]
    g.step(SUB, 1)
    st = g.read_state()
    ref = _state_from_dump("ring_cap", "sub1")
    # the particle at x = 1e7 (hash products beyond 2^53) is culled like the reference
    assert np.array_equal(st["active"], ref["active"])


def test_deterministic_repeat():
    """Two runs from the same inputs are bit-identical (gather-only kernels,
    stable grid order)."""
    log("case determinism single_gpu")
    from fluidphysics_amd import from_scene, scenes
    sc = scenes.single_gpu_scene()
    outs = []
    for _ in range(2):
        g = from_scene(sc)
        for _f in range(5):
            g.step(sc.dt, sc.substeps)
        outs.append(g.read_state())
        g.close()
    for k in ("pos", "vel", "active"):
        assert np.array_equal(outs[0][k], outs[1][k])


@pytest.mark.gpu
def test_pair_list_overflow_regathers_bit_identically():
    """Particles whose pairs do not fit the recorded list are gathered again
    from the grid by the displacement and XSPH passes: a run with a tiny
    list (most particles overflow) equals the default run bit for bit."""
    log("case pair-list overflow single_gpu, pair_slots 8 vs default")
    from fluidphysics_amd import from_scene, scenes
    sc = scenes.single_gpu_scene()
    outs, ovf = [], []
    for slots in (0, 8):
        g = from_scene(sc, pair_slots=slots)
        for _f in range(4):
            g.step(sc.dt, sc.substeps)
        outs.append(g.read_state())
        ovf.append(g.info().pair_overflow)
        g.close()
    log("  pair_overflow default %d, pair_slots 8 %d" % (ovf[0], ovf[1]))
    assert ovf[1] > 0
    for k in ("pos", "prev", "vel", "active"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_pass_profile_names():
    """Per-pass event timing reports every pass of a substep under its name."""
    log("case pass profile plain_cpu")
    d = G.scenario("plain_cpu")
    g = make_gpu(d)
    g.profile(True)
    g.step(1.0 / 60.0, d["substeps"])
    g.sync()
    stats = g.profile_read()
    g.profile(False)
    # bounds and velocity run inside the last constraint iteration (solids)
    # and the displacement (fluid): no pass of their own with constraints
    for name in ("integrate", "contacts", "resort", "density", "displace", "constraints", "xsph"):
        assert name in stats and stats[name][0] == d["substeps"], name
    g.close()


def test_grab_selection_matches_the_callers_rule():
    """fp_select_solids against src/main.js:293-301 evaluated on the host:
    solids of the picked solid's component within 0.35 of the hit point, the
    distance in double from the float32 positions."""
    from fluidphysics_amd import from_scene, scenes
    sc = scenes.single_gpu_scene(frames=40)
    log("case grab selection %s" % sc.name)
    s = from_scene(sc)
    cx, cy, cz = sc.cube[:3]
    for f in range(30):
        if f == 5:
            # split the cube in two: the plane x = cx + 0.013 (between two
            # lattice layers), a wedge wider than the cube
            d1 = np.array([0.0, -1.0, -0.9]) / np.linalg.norm([0.0, -1.0, -0.9])
            d2 = np.array([0.0, -1.0, 0.9]) / np.linalg.norm([0.0, -1.0, 0.9])
            assert s.cut(cx + 0.013, cy + 5.0, cz, *d1, *d2) > 0
        s.step(1.0 / 60.0, 5)
    st = s.read_state()
    pos = st["pos"].reshape(-1, 3)[: s.solidCount].astype(np.float64)
    group = st["group"]
    assert len(np.unique(group)) > 1
    rng = np.random.default_rng(3)
    checked = 0
    for trial in range(12):
        o = np.array([cx, cy, cz]) + rng.uniform(-2, 2, 3) + np.array([0.0, 2.0, 3.0])
        target = np.array([cx, cy, cz]) + rng.uniform(-0.4, 0.4, 3)
        d = (target - o) / np.linalg.norm(target - o)
        hit = s.pickSolid(*o, *d, 0.25)
        if hit is None:
            continue
        hp = o + d * hit["t"]
        ids = s.select_solids(hp, 0.35, hit["id"])
        dd = pos - hp
        want = np.flatnonzero((group == group[hit["id"]]) & ((dd * dd).sum(axis=1) < 0.35 * 0.35))
        assert np.array_equal(ids, want), (trial, len(ids), len(want))
        every = s.select_solids(hp, 0.35, -1)
        assert np.array_equal(every, np.flatnonzero((dd * dd).sum(axis=1) < 0.35 * 0.35))
        checked += 1
    assert checked >= 6
    s.close()


def _split_scene_solver(frames_before_cut=5):
    """A single-GPU scene whose cube a planar wedge splits in two (the plane
    x = cx + 0.013, between two lattice layers), stepped to the cut."""
    from fluidphysics_amd import from_scene, scenes
    sc = scenes.single_gpu_scene(frames=40)
    s = from_scene(sc)
    for _ in range(frames_before_cut):
        s.step(1.0 / 60.0, 5)
    cx, cy, cz = sc.cube[:3]
    d1 = np.array([0.0, -1.0, -0.9]) / np.linalg.norm([0.0, -1.0, -0.9])
    d2 = np.array([0.0, -1.0, 0.9]) / np.linalg.norm([0.0, -1.0, 0.9])
    return sc, s, (cx + 0.013, cy + 5.0, cz, *d1, *d2)


def _csgraph_partition(ns, c):
    """Connected components of the solids over the alive constraints, by
    scipy's textbook routine (the check is independent of both union-finds)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    keep = c["alive"] != 0
    g = coo_matrix((np.ones(int(keep.sum())), (c["a"][keep], c["b"][keep])), shape=(ns, ns))
    return connected_components(g, directed=False)


def _same_partition(x, y):
    """x and y label the same partition: the pairs (x_i, y_i) form a bijection."""
    pairs = np.unique(np.stack([x, y], axis=1), axis=0)
    return len(pairs) == len(np.unique(x)) == len(np.unique(y))


def test_components_on_device_match_the_partition():
    """fp_components (the device form of solver.js:488-499) after a cut that
    splits the cube and a grab that tears it: the partition equals scipy's
    connected components of the alive constraints and the host union-find's
    (whose roots are the reference's labels); every label is its component's
    smallest solid index."""
    sc, s, knife = _split_scene_solver()
    log("case components: %s" % sc.name)
    labels0, n0 = s.components()
    assert n0 == 1 and not labels0.any()  # the lattice is one body
    assert s.cut(*knife) > 0
    cx, cy, cz = sc.cube[:3]
    s.grabs = [{"i": i, "x": cx - 3.0, "y": cy + 3.0, "z": cz} for i in range(24)]
    for _ in range(12):
        s.step(1.0 / 60.0, 5)
    s.grabs = []
    s.step(1.0 / 60.0, 5)
    ns = s.solidCount
    labels, n = s.components()
    c = s.read_constraints()
    ncc, ref = _csgraph_partition(ns, c)
    group = s.read_state()["group"][:ns]
    log("components %d (scipy %d), constraints alive %d of %d" % (n, ncc, int(c["alive"].sum()), len(c["alive"])))
    assert n >= 2 and n == ncc == len(np.unique(labels))
    assert _same_partition(labels, ref) and _same_partition(labels, group)
    for lab in np.unique(labels):
        assert lab == np.flatnonzero(labels == lab)[0]
    s.close()


def test_selection_after_a_cut_sees_the_components_of_before_it():
    """Between a cut and the next step the caller's group still holds the
    labels of before the cut (solver.js relabels at a step's end), so the
    grab selection takes the whole uncut body; after the step, one side."""
    sc, s, knife = _split_scene_solver()
    log("case selection across a cut: %s" % sc.name)
    cx, cy, cz = sc.cube[:3]
    assert s.cut(*knife) > 0
    st = s.read_state()
    ns = s.solidCount
    pos = st["pos"].reshape(-1, 3)[:ns].astype(np.float64)
    group = st["group"][:ns]
    assert len(np.unique(group)) == 1  # not relabelled yet
    hp = pos.mean(axis=0)
    hp[0] = cx + 0.013  # on the cut plane, at the body's centre
    d2 = ((pos - hp) ** 2).sum(axis=1)
    near = np.flatnonzero(d2 < 0.35 * 0.35)
    assert (pos[near, 0] < hp[0]).any() and (pos[near, 0] > hp[0]).any()
    pick = int(near[0])
    assert np.array_equal(s.select_solids(hp, 0.35, pick), near)
    s.step(1.0 / 60.0, 5)
    st = s.read_state()
    pos = st["pos"].reshape(-1, 3)[:ns].astype(np.float64)
    group = st["group"][:ns]
    assert len(np.unique(group)) >= 2
    d2 = ((pos - hp) ** 2).sum(axis=1)
    want = np.flatnonzero((group == group[pick]) & (d2 < 0.35 * 0.35))
    got = s.select_solids(hp, 0.35, pick)
    assert np.array_equal(got, want) and len(got) < len(np.flatnonzero(d2 < 0.35 * 0.35))
    s.close()
