"""Pins of the oracle's Jacobi formulation and tilted gravity (CPU only).

The HIP kernels are checked against the oracle's Jacobi mode at scale, and
the reference (src/solver.js) has no Jacobi mode and no tilted gravity, so
these functions are pinned here by what the mathematics fixes:

  * on one isolated constraint or one isolated contact the Jacobi gather
    and the reference's Gauss-Seidel sweep (solver.js:246-279, :356-381) do
    the same arithmetic, so they must agree bit for bit (with a relaxation
    omega / max(deg_a, deg_b) of exactly 1);
  * an XPBD projection leaves the residual C * alpha / (w + alpha)
    (solver.js:372: dl = -C / (w + alpha)); a contact projection leaves the
    pair exactly at rsum;
  * tearing (d > rest * breakStrain, :369) and plasticity (|C / rest| >
    plasticThreshold, :378) decide identically in both modes exactly at and
    one float32 step past their thresholds;
  * momentum: every Jacobi constraint iteration and every Jacobi contact
    pass moves sum(m * dp) by zero (each pair's two shares are equal and
    opposite -- which needs the symmetric scale max(deg_a, deg_b) and both
    endpoints seeing the pair);
  * tilted gravity: a free particle follows the closed form of integrate +
    velocity damping (solver.js:194-202, :219-234) along the tilted vector.

A sign error, a dropped mass factor, an asymmetric scale or a gravity
component on the wrong axis fails at least one of these.
"""
import numpy as np
import pytest

from oracle.oracle import OracleSolver

H_SUB = 1.0 / 300.0


def _cube2(storage, mode, **kw):
    """A 2x2x2 lattice (every corner bonded to the other 7: static degree 7)."""
    p = dict(gravity=0.0, constraint_iters=1, jacobi_omega=7.0, compliance=2.5e-4)
    p.update(kw)
    o = OracleSolver(p, 0, storage, mode)
    o.add_solid_cube(0.0, 2.0, 0.0, 2, 0.18, 0.08)
    return o


def _one_constraint(o, p0, p1, w=(12.5, 20.0), rest=None):
    """Only constraint (0, 1) alive; particles 0 and 1 at p0, p1; the other
    six far apart with no alive constraint."""
    c = o.read_constraints()
    assert c["a"][0] == 0 and c["b"][0] == 1
    alive = np.zeros_like(c["alive"])
    alive[0] = 1
    r = c["rest"].copy()
    if rest is not None:
        r[0] = rest
    o.write_constraints(r, alive)
    st = o.read_state()
    pos = st["pos"].reshape(-1, 3).copy()
    pos[0], pos[1] = p0, p1
    for k in range(2, 8):
        pos[k] = (3.0 * k, 2.0, -3.0 * k)
    im = st["inv_mass"].copy()
    im[0], im[1] = w
    o.write_state(pos, pos, np.zeros_like(pos), im, st["active"])


def _pos(o):
    return o.read_state64()["pos"].reshape(-1, 3) if o.storage == "f64" else \
        o.read_state()["pos"].reshape(-1, 3).astype(np.float64)


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_single_constraint_jacobi_equals_gauss_seidel(storage):
    p0, p1 = (0.0, 2.0, 0.0), (0.3, 2.1, 0.05)
    outs = {}
    for mode in ("gs", "jacobi"):
        o = _cube2(storage, mode)
        _one_constraint(o, p0, p1)
        o.solve_constraints(H_SUB, 1)
        outs[mode] = (_pos(o), o.read_constraints())
    assert np.array_equal(outs["gs"][0], outs["jacobi"][0])
    assert np.array_equal(outs["gs"][1]["rest"], outs["jacobi"][1]["rest"])
    assert np.array_equal(outs["gs"][1]["alive"], outs["jacobi"][1]["alive"])
    moved = outs["gs"][0][:2] - np.array([p0, p1], np.float32).astype(np.float64)
    assert np.abs(moved).max() > 1e-3  # the constraint was stretched and projected


def test_projection_leaves_the_compliant_residual():
    """|p0 - p1| - rest after one projection = C * alpha / (w + alpha), and
    the two ends move in inverse proportion to their masses."""
    # positions enter the oracle as float32 (the reference's Float32Array)
    p0 = np.array([0.0, 2.0, 0.0], np.float32).astype(np.float64)
    p1 = np.array([0.3, 2.1, 0.05], np.float32).astype(np.float64)
    for mode in ("gs", "jacobi"):
        o = _cube2("f64", mode, plastic_threshold=10.0)  # no plasticity
        _one_constraint(o, p0, p1)
        rest = o.read_constraints()["rest"][0]
        o.solve_constraints(H_SUB, 1)
        q = _pos(o)
        wa, wb = 12.5, 20.0
        alpha = 2.5e-4 / H_SUB ** 2
        c0 = np.linalg.norm(p0 - p1) - rest
        c1 = np.linalg.norm(q[0] - q[1]) - rest
        assert c1 == pytest.approx(c0 * alpha / (wa + wb + alpha), rel=1e-12, abs=1e-15)
        da, db = q[0] - p0, q[1] - p1
        assert np.allclose(da / wa, -db / wb, rtol=1e-12, atol=1e-16)  # m_a da = -m_b db


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_tear_and_plastic_decided_at_their_thresholds(storage):
    """rest = 1, breakStrain = 2.5, plasticThreshold = 0.25 (all exact): a
    pair at d = 2.5 is kept, one float32 step farther it tears; a strain of
    exactly 0.25 keeps rest, one float32 step more adds C * plasticRate."""
    up = float(np.nextafter(np.float32(2.5), np.float32(3)))
    up_p = float(np.nextafter(np.float32(1.25), np.float32(2)))
    # (x, plastic threshold, alive after, rest after; None = 1 + C * rate)
    cases = [(2.5, 10.0, True, 1.0), (up, 10.0, False, 1.0), (1.25, 0.25, True, 1.0), (up_p, 0.25, True, None)]
    for x, thr, alive_after, rest_after in cases:
        got = []
        for mode in ("gs", "jacobi"):
            o = _cube2(storage, mode, break_strain=2.5, plastic_threshold=thr, plastic_rate=0.04)
            _one_constraint(o, (0.0, 2.0, 0.0), (x, 2.0, 0.0), rest=1.0)
            o.solve_constraints(H_SUB, 1)
            c = o.read_constraints()
            got.append((bool(c["alive"][0]), float(c["rest"][0]), _pos(o)[:2].copy()))
            assert got[-1][0] == alive_after, (mode, x)
            if rest_after is not None:
                assert got[-1][1] == rest_after, (mode, x)
            else:
                assert got[-1][1] == 1.0 + (x - 1.0) * 0.04, (mode, x)
            if not alive_after:  # a torn constraint moves nothing
                assert np.array_equal(got[-1][2], np.array([[0, 2, 0], [x, 2, 0]], np.float64))
        assert got[0][0] == got[1][0] and got[0][1] == got[1][1]
        assert np.array_equal(got[0][2], got[1][2])


def _jittered_cube(n, seed, mode, storage="f64", kill=0.2, **kw):
    p = dict(gravity=0.0, constraint_iters=1, jacobi_omega=1.5)
    p.update(kw)
    o = OracleSolver(p, 512, storage, mode)
    o.add_solid_cube(0.0, 2.0, 0.0, n, 0.18, 0.08)
    rng = np.random.default_rng(seed)
    st = o.read_state()
    ns = n ** 3
    pos = st["pos"].reshape(-1, 3).copy()
    pos[:ns] += rng.uniform(-0.03, 0.03, (ns, 3))
    im = st["inv_mass"].copy()
    im[:ns] = rng.uniform(5.0, 20.0, ns).astype(np.float32)
    o.write_state(pos, pos, np.zeros_like(pos), im, st["active"])
    c = o.read_constraints()
    alive = (rng.uniform(0, 1, len(c["alive"])) >= kill).astype(np.uint8)
    o.write_constraints(None, alive)
    return o, rng


def _momentum_change(o, before):
    st = o.read_state()
    act = st["active"].astype(bool)
    m = np.zeros(len(act))
    m[act] = 1.0 / st["inv_mass"][act].astype(np.float64)
    dp = o.read_state64()["pos"].reshape(-1, 3) - before
    mdp = m[:, None] * dp
    return mdp.sum(axis=0), np.abs(mdp).sum()


def test_jacobi_constraint_iteration_conserves_momentum():
    o, _ = _jittered_cube(5, 7, "jacobi")
    before = o.read_state64()["pos"].reshape(-1, 3).copy()
    o.solve_constraints(H_SUB, 1)
    tot, mag = _momentum_change(o, before)
    assert mag > 1e-4
    assert np.abs(tot).max() <= 1e-12 * mag


def test_jacobi_contacts_conserve_momentum():
    """Solids packed closer than their contact distance (every bond dead, so
    every close pair collides) with fluid scattered through them."""
    o, rng = _jittered_cube(5, 9, "jacobi", kill=1.1)
    st = o.read_state()
    pos = st["pos"].reshape(-1, 3)
    sol = pos[:125].astype(np.float64)
    # squeeze the lattice to 0.12 spacing about its centre: solid pairs overlap
    ctr = sol.mean(axis=0)
    sol = ctr + (sol - ctr) * (0.12 / 0.18)
    pos = pos.copy()
    pos[:125] = sol
    o.write_state(pos, pos, None, None, None)
    fl = ctr + rng.uniform(-0.3, 0.3, (200, 3))
    o.emit(fl, np.zeros_like(fl), 0.03)
    # inverse masses as float32 values, so that m = 1 / w below is the
    # oracle's own mass to double precision
    o.write_state(None, None, None, o.read_state()["inv_mass"], None)
    before = o.read_state64()["pos"].reshape(-1, 3).copy()
    o.solve_contacts()
    tot, mag = _momentum_change(o, before)
    moved = np.abs(o.read_state64()["pos"].reshape(-1, 3) - before).max(axis=1)
    assert (moved[:125] > 0).sum() > 50 and (moved[125:] > 0).sum() > 20
    assert np.abs(tot).max() <= 1e-12 * mag


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_single_contact_jacobi_equals_gauss_seidel(storage):
    """One solid and one fluid particle inside rs + rf: the Jacobi gather
    (contact_scale 1) and the reference's sweep give the same bits, and the
    pair ends exactly rsum apart."""
    outs = []
    for mode in ("gs", "jacobi"):
        # a table large enough that the 27 probes hit distinct buckets (the
        # reference's sweep visits a repeated bucket twice)
        o = OracleSolver(dict(gravity=0.0, contact_scale=1.0, table_size=1 << 20), 4, storage, mode)
        o.add_solid_cube(0.0, 2.0, 0.0, 1, 0.18, 0.08)
        o.emit(np.array([[0.1, 2.05, 0.02]]), None, 0.03)
        o.solve_contacts()
        outs.append(_pos(o)[:2].copy())
    assert np.array_equal(outs[0], outs[1])
    d = np.linalg.norm(outs[0][0] - outs[0][1])
    assert d == pytest.approx(0.09 + 0.075, rel=1e-12 if storage == "f64" else 1e-6)


@pytest.mark.parametrize("phi", [0.0, 1.1, -2.5])
def test_tilted_gravity_free_fall_closed_form(phi):
    """A fluid particle far from everything under g = 9.8 * (sin t cos phi,
    -cos t, sin t sin phi), t = 15 deg, for K substeps: v_{k+1} =
    damp * (v_k + g h) with damp = 1 - 0.6 h, so the displacement is
    g h^2 * sum_{k<K} (1 - damp^(k+1)) / (1 - damp)."""
    t = np.radians(15.0)
    g = 9.8 * np.array([np.sin(t) * np.cos(phi), -np.cos(t), np.sin(t) * np.sin(phi)])
    o = OracleSolver(dict(gravity=g[1], gravity_x=g[0], gravity_z=g[2]), 4, "f64", "jacobi")
    o.add_solid_cube(-2.5, 0.09, -2.5, 1, 0.18, 0.08)
    x0 = np.array([0.5, 2.0, 0.3])
    o.emit(x0[None, :], None, 0.03)
    K = 12
    for _ in range(K):
        o.substep(H_SUB)
    ns = 1
    x = o.read_state64()["pos"].reshape(-1, 3)[ns]
    damp = 1.0 - 0.6 * H_SUB
    s = sum((1.0 - damp ** (k + 1)) / (1.0 - damp) for k in range(K))
    want = x0 + g * H_SUB ** 2 * s
    assert np.allclose(x, want, rtol=0, atol=1e-13)
    v = o.read_state64()["vel"].reshape(-1, 3)[ns]
    assert np.allclose(v, g * H_SUB * damp * (1.0 - damp ** K) / (1.0 - damp), rtol=1e-10, atol=1e-14)


# ---- the GPU twin's optional solid modes: volume tets, solid contact cache ----

def _tet_cube(mode_kw, n=2, seed=3, jitter=0.02, masses=True):
    p = dict(gravity=0.0, constraint_iters=1, jacobi_omega=1.0, table_size=1 << 16)
    p.update(mode_kw)
    o = OracleSolver(p, 0, "f64", "jacobi")
    o.add_solid_cube(0.0, 2.0, 0.0, n, 0.18, 0.08)
    rng = np.random.default_rng(seed)
    st = o.read_state()
    ns = n ** 3
    pos = st["pos"].reshape(-1, 3).copy()
    pos[:ns] += rng.uniform(-jitter, jitter, (ns, 3))
    im = st["inv_mass"].copy()
    if masses:
        im[:ns] = rng.uniform(5.0, 20.0, ns).astype(np.float32)
    o.write_state(pos, pos, np.zeros_like(pos), im, st["active"])
    return o


def _signed_volume(q):
    return np.dot(np.cross(q[1] - q[0], q[2] - q[0]), q[3] - q[0]) / 6.0


def test_lattice_tets_tile_the_cube_with_positive_rest_volumes():
    """Five tets per cell (the twin's split, mirrored on parity): every rest
    volume positive and equal to the signed volume of its vertices, and each
    cell's five add up to the cell's volume."""
    o = _tet_cube(dict(tet_volume=1), n=4, jitter=0.0, masses=False)
    t = o.read_tets()
    pos = o.read_state64()["pos"].reshape(-1, 3)
    assert len(t["v0"]) == 5 * 27 and (t["v0"] > 0).all() and t["alive"].all()
    vols = np.array([_signed_volume(pos[v]) for v in t["v"]])
    # (the lattice entered as float32 positions after V0 was taken in double)
    assert np.allclose(vols, t["v0"], rtol=1e-5, atol=0)
    assert np.allclose(t["v0"].reshape(27, 5).sum(axis=1), 0.18 ** 3, rtol=1e-6)


def test_tet_step_is_the_volume_gradient_projection():
    """One tet alive in a jittered 2^3 cell: each vertex moves by
    -C s w_i dV/dp_i / sum_q w_q |dV/dp_q|^2 with the gradient taken by
    central differences of the signed volume (independent of the cross
    products), C = V - V0, s = 1 / max(incident tets) = 1/4; momentum
    conserved; the volume error shrinks to (1 - s) C up to second order."""
    o = _tet_cube(dict(tet_volume=1))
    t = o.read_tets()
    alive = np.zeros_like(t["alive"])
    alive[0] = 1
    o.write_tet_alive(alive)
    st = o.read_state()
    w = st["inv_mass"].astype(np.float64)
    x = o.read_state64()["pos"].reshape(-1, 3).copy()
    v = t["v"][0]
    q = x[v]
    C = _signed_volume(q) - t["v0"][0]
    g = np.zeros((4, 3))
    eps = 1e-6
    for m in range(4):
        for c in range(3):
            qp, qm = q.copy(), q.copy()
            qp[m, c] += eps
            qm[m, c] -= eps
            g[m, c] = (_signed_volume(qp) - _signed_volume(qm)) / (2 * eps)
    denom = sum(w[v[m]] * g[m] @ g[m] for m in range(4))
    s = 0.25
    o.solid_extra_step()
    y = o.read_state64()["pos"].reshape(-1, 3)
    for m in range(4):
        want = -C * s * w[v[m]] * g[m] / denom
        assert np.allclose(y[v[m]] - x[v[m]], want, rtol=1e-6, atol=1e-15), m
    others = np.setdiff1d(np.arange(8), v)
    assert np.array_equal(y[others], x[others])
    dp = y[v] - x[v]
    assert np.abs(((1.0 / w[v])[:, None] * dp).sum(axis=0)).max() < 1e-15
    c1 = _signed_volume(y[v]) - t["v0"][0]
    assert abs(c1 - (1 - s) * C) < 50 * C * C / t["v0"][0] + 1e-18
    assert abs(C) > 1e-6  # the jitter made a real volume error


def test_tet_step_conserves_momentum_with_every_tet():
    o = _tet_cube(dict(tet_volume=1), n=4, seed=5)
    st = o.read_state()
    m = 1.0 / st["inv_mass"].astype(np.float64)
    x = o.read_state64()["pos"].reshape(-1, 3).copy()
    o.solid_extra_step()
    dp = o.read_state64()["pos"].reshape(-1, 3) - x
    mdp = m[:, None] * dp
    assert np.abs(mdp).sum() > 1e-6
    assert np.abs(mdp.sum(axis=0)).max() <= 1e-12 * np.abs(mdp).sum()


def _two_solids(d, adjacent, rs=0.09, wi=12.5, wj=20.0):
    """Solids 0 and (1 if adjacent else 2) of a 3^3 lattice (spacing 0.18)
    placed d apart along x, every other solid far away, every bond dead."""
    p = dict(gravity=0.0, constraint_iters=1, jacobi_omega=1.0, table_size=1 << 16, solid_radius=rs,
             solid_contact_cache=1)
    o = OracleSolver(p, 0, "f64", "jacobi")
    o.add_solid_cube(0.0, 2.0, 0.0, 3, 0.18, 0.08)
    o.write_constraints(None, np.zeros(o.info()["num_constraints"], np.uint8))
    j = 9 if adjacent else 18  # (1,0,0) or (2,0,0)
    st = o.read_state()
    pos = st["pos"].reshape(-1, 3).copy()
    for k in range(27):
        pos[k] = (3.0 + 0.5 * (k % 5), 2.0 + 0.5 * (k // 5), -3.0)
    pos[0] = (0.0, 2.0, 0.0)
    pos[j] = (np.float32(d), 2.0, 0.0)
    im = st["inv_mass"].copy()
    im[0], im[j] = wi, wj
    o.write_state(pos, pos, np.zeros_like(pos), im, st["active"])
    return o, j


@pytest.mark.parametrize("adjacent", [False, True])
def test_cached_solid_contact_halves_the_overlap(adjacent):
    """Two unbonded solids inside their contact distance: the cache lists
    each for the other; one extra step closes half the overlap (the 0.5 of
    wgsl.js:594) with equal and opposite momenta.  Lattice neighbours (a cut
    face) use min(2 r, 0.97 spacing) as their contact distance."""
    d0 = float(np.float32(0.15))
    o, j = _two_solids(d0, adjacent)
    o.collect_contacts()
    cache = o.read_cache()
    assert list(cache["ids"][0][: cache["n"][0]]) == [j] and list(cache["ids"][j][: cache["n"][j]]) == [0]
    rsum = min(0.18, 0.97 * 0.18) if adjacent else 0.18
    x = o.read_state64()["pos"].reshape(-1, 3).copy()
    o.solid_extra_step()
    y = o.read_state64()["pos"].reshape(-1, 3)
    d1 = np.linalg.norm(y[j] - y[0])
    assert d1 == pytest.approx(d0 + 0.5 * (rsum - d0), rel=1e-12)
    assert np.allclose((y[0] - x[0]) / 12.5, -(y[j] - x[j]) / 20.0, rtol=1e-12, atol=1e-18)


def test_cached_contact_correction_is_capped():
    """A deep overlap moves each solid by at most half a solid radius per
    iteration (applySolidContacts, wgsl.js:598-606)."""
    o, j = _two_solids(float(np.float32(0.01)), False, wi=20.0, wj=12.5)
    o.collect_contacts()
    x = o.read_state64()["pos"].reshape(-1, 3).copy()
    o.solid_extra_step()
    y = o.read_state64()["pos"].reshape(-1, 3)
    for k in (0, j):
        assert np.linalg.norm(y[k] - x[k]) <= 0.5 * 0.09 * (1 + 1e-12)
    assert np.linalg.norm(y[0] - x[0]) == pytest.approx(0.045, rel=1e-12)


def test_cache_keeps_the_smallest_partners_and_counts_the_rest():
    """More unbonded partners in range than the cache holds (16): the 16
    smallest indices are kept, in order, and the rest are counted."""
    p = dict(gravity=0.0, constraint_iters=1, table_size=1 << 16, solid_contact_cache=1, solid_radius=0.09,
             cell_size=0.4)
    o = OracleSolver(p, 0, "f64", "jacobi")
    o.add_solid_cube(0.0, 2.0, 0.0, 3, 0.18, 0.08)
    o.write_constraints(None, np.zeros(o.info()["num_constraints"], np.uint8))
    st = o.read_state()
    rng = np.random.default_rng(8)
    pos = 0.0 + rng.uniform(-0.04, 0.04, (27, 3)) + np.array([0.0, 2.0, 0.0])
    o.write_state(pos, pos, None, None, None)
    o.collect_contacts()
    c = o.read_cache()
    assert c["n"][0] == 16 and list(c["ids"][0]) == list(range(1, 17))
    assert c["overflow"] >= 27 * 10


def test_tet_tears_at_its_threshold():
    """A tet edge longer than its lattice rest + tear_gap * 2 r kills the tet;
    exactly at it the tet lives (wgsl.js:478-486)."""
    for extra, alive in ((0.0, True), (1e-6, False)):
        o = _tet_cube(dict(tet_volume=1, tear_gap=1.0), jitter=0.0, masses=False)
        t = o.read_tets()
        v = t["v"][0]
        pos = o.read_state64()["pos"].reshape(-1, 3).copy()
        a, b = v[0], v[1]
        # move b along a-b until |a - b| = rest + gap (+ extra)
        rest = np.linalg.norm(pos[b] - pos[a])
        target = rest + 2 * 0.09 + extra
        u = (pos[b] - pos[a]) / rest
        pos[b] = pos[a] + u * target
        o.write_state(pos.astype(np.float32), None, None, None, None)
        q = o.read_state64()["pos"].reshape(-1, 3)
        dist = np.linalg.norm(q[b] - q[a])
        o.tet_update()
        got = bool(o.read_tets()["alive"][0])
        assert got == (dist <= rest + 2 * 0.09), (extra, dist, rest + 0.18)
    assert not got


def test_cut_kills_the_tets_whose_edges_it_cuts():
    """After a cut, a tet is dead exactly when one of its six edges -- every
    one a lattice constraint -- was cut."""
    o = _tet_cube(dict(tet_volume=1), n=4, jitter=0.005, masses=False)
    cons = o.read_constraints()
    edge = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(cons["a"], cons["b"]))}
    # the plane x = 0.01, between two lattice layers, across the whole cube
    cam = np.array([0.01, 3.5, 0.0])
    d1 = np.array([0.0, -1.0, -0.9])
    d2 = np.array([0.0, -1.0, 0.9])
    d1 /= np.linalg.norm(d1)
    d2 /= np.linalg.norm(d2)
    assert o.cut(cam, d1, d2) > 0
    alive_c = o.read_constraints()["alive"]
    t = o.read_tets()
    for v, al in zip(t["v"], t["alive"]):
        cut = any(not alive_c[edge[(min(p, q), max(p, q))]] for i, p in enumerate(v) for q in v[i + 1:])
        assert bool(al) == (not cut)
    assert 0 < int(t["alive"].sum()) < len(t["alive"])
