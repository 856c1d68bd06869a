"""Loading the golden vectors of tests/golden/ and replaying their scenarios.

The vectors were written by the reference solver (src/solver.js) under node
(tests/golden/make_golden.mjs); the scenario inputs come from
tests/golden/gen_inputs.py, which this module calls again so the inputs need
not be stored.
"""
import gzip
import hashlib
import importlib.util
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _gen_module():
    spec = importlib.util.spec_from_file_location("gen_inputs", os.path.join(GOLDEN, "gen_inputs.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def scenario(name):
    return getattr(_gen_module(), name)()


def trace(name):
    with open(os.path.join(GOLDEN, name + "_trace.json")) as f:
        return json.load(f)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load_grid0(name, table_size, capacity):
    with gzip.open(os.path.join(GOLDEN, name + "_grid0.bin.gz")) as f:
        raw = f.read()
    cs = np.frombuffer(raw, np.int32, table_size + 1, 0)
    ce = np.frombuffer(raw, np.int32, capacity, 4 * (table_size + 1))
    return cs, ce


def load_state(name, suffix, capacity, solid_count, num_cons):
    """State dump layout: pos f32[3N] prev f32[3N] vel f32[3N] rho f32[N]
    rhoNear f32[N] group i32[Ns] consRest f64[C] active u8[N] consAlive u8[C]."""
    with gzip.open(os.path.join(GOLDEN, "%s_%s.bin.gz" % (name, suffix))) as f:
        raw = f.read()
    out, off = {}, 0
    for key, dt, n in (("pos", np.float32, 3 * capacity), ("prev", np.float32, 3 * capacity),
                       ("vel", np.float32, 3 * capacity), ("rho", np.float32, capacity),
                       ("rho_near", np.float32, capacity), ("group", np.int32, solid_count),
                       ("rest", np.float64, num_cons), ("active", np.uint8, capacity),
                       ("alive", np.uint8, num_cons)):
        out[key] = np.frombuffer(raw, dt, n, off).copy()
        off += np.dtype(dt).itemsize * n
    assert off == len(raw)
    return out


def make_oracle(d, storage="f32", mode="gs", iters=1, **overrides):
    from oracle.oracle import OracleSolver
    params = dict(gravity=d["gravity"], constraint_iters=iters)
    params.update(overrides)
    s = OracleSolver(params, d["maxFluid"], storage=storage, mode=mode)
    s.add_solid_cube(*d["cube"])
    e = np.asarray(d["emit"], dtype=np.float64)
    if len(e):
        s.emit(e[:, 0:3], e[:, 3:6], e[:, 6])
    return s


class Driver:
    """The caller logic of make_golden.mjs (cuts, pick + grab + drag), applied
    to any solver with the OracleSolver / fluidphysics_amd.Solver surface."""

    def __init__(self, d):
        self.d = d
        self.grab_targets = None
        self.cuts = []
        self.picks = []

    def before_frame(self, s, f, state_pos, group):
        d = self.d
        for c in d["cuts"]:
            if c["frame"] == f:
                self.cuts.append((f, s.cut(c["o"], c["d1"], c["d2"])))
        g = d.get("grab")
        if not g:
            return
        if f == g["frame"]:
            hit = s.pick_solid(g["o"], g["d"], g["maxPerp"])
            self.picks.append((f, hit))
            if hit is not None:
                hid, t = hit
                hp = [g["o"][k] + g["d"][k] * t for k in range(3)]
                p = state_pos().reshape(-1, 3)
                grp = group()
                self.grab_targets = []
                for i in range(len(grp)):
                    if grp[i] != grp[hid]:
                        continue
                    dx = float(p[i, 0]) - hp[0]
                    dy = float(p[i, 1]) - hp[1]
                    dz = float(p[i, 2]) - hp[2]
                    if dx * dx + dy * dy + dz * dz < g["radius"] * g["radius"]:
                        self.grab_targets.append([i, hp[0] + dx, hp[1] + dy, hp[2] + dz])
        if self.grab_targets is not None and g["frame"] < f < g["until"]:
            for t in self.grab_targets:
                t[1] += g["drag"][0]
                t[2] += g["drag"][1]
                t[3] += g["drag"][2]
        if self.grab_targets is not None and f == g["until"]:
            self.grab_targets = []
        tg = self.grab_targets or []
        s.set_grabs(np.array([t[0] for t in tg], np.int32),
                    np.array([t[1:] for t in tg], np.float64).reshape(-1, 3))


def frame_stats(st, solid_count, rho=None):
    """Kinetic energy, mean density, fluid count (make_golden.mjs frameStats)."""
    act = st["active"].astype(bool)
    v = st["vel"].astype(np.float64).reshape(-1, 3)[act]
    m = 1.0 / st["inv_mass"].astype(np.float64)[act]
    ke = float(np.sum(0.5 * m * np.sum(v * v, axis=1)))
    fl = act.copy()
    fl[:solid_count] = False
    mean_rho = float(np.mean(rho[fl])) if (rho is not None and fl.any()) else 0.0
    return dict(ke=ke, meanRho=mean_rho, fluid=int(fl.sum()))
