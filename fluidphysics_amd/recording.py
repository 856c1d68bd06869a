"""Recording and replay of a scene session.

A recording is the scene's state blob at the start (Solver.save_state) plus
the caller-side inputs that followed, each stamped with the frame it came
before: emits, slices, grab lists, resets, tuning changes and the steps
themselves.  Replaying feeds the same inputs at the same frames into a
solver restored from the blob; the GPU path is deterministic, so the replay
ends in the same bits.  The reference's GPU page records the same way -- a
snapshot when recording starts, then the input log, replayed at the same
frame indices (src/gpu/main.js:308-390, snapshot / restore in
src/gpu/sim.js:494-547).

    rec = Recorder(solver, setup=dict(max_fluid=..., cube=(...), params={...}))
    rec.cut(...); rec.grabs([...]); rec.step(1 / 60, 5) ...
    bundle = rec.bundle()            # JSON-serializable
    replayed = replay(bundle)        # a new Solver at the recording's end
"""
import base64

import numpy as np

from . import Solver

BUNDLE_TYPE = "fluidphysics-recording"


class Recorder:
    """Drives `solver` and logs each input.  `setup` describes how to build a
    solver of the same shape: {'max_fluid', 'cube': (cx, cy, cz, n, spacing,
    mass), 'params': {...}} -- a replay rebuilds it and then loads the state
    blob taken here."""

    def __init__(self, solver, setup):
        self.s = solver
        self.setup = {"max_fluid": int(setup["max_fluid"]), "cube": [float(v) for v in setup["cube"]],
                      "params": {k: (float(v) if isinstance(v, float) else v) for k, v in setup["params"].items()}}
        self.frame = 0
        self.actions = []
        self.state = solver.save_state()

    def _log(self, op, **args):
        self.actions.append({"f": self.frame, "op": op, **args})

    def emit_batch(self, pos, vel, mass=0.03):
        pos = np.asarray(pos, np.float64).reshape(-1, 3)
        vel = np.asarray(vel, np.float64).reshape(-1, 3)
        m = np.broadcast_to(np.asarray(mass, np.float64), (len(pos),))
        self._log("emit", pos=pos.ravel().tolist(), vel=vel.ravel().tolist(), mass=m.tolist())
        self.s.emit_batch(pos, vel, m)

    def cut(self, o, d1, d2):
        self._log("cut", o=list(map(float, o)), d1=list(map(float, d1)), d2=list(map(float, d2)))
        return self.s.cut(*o, *d1, *d2)

    def grabs(self, grabs):
        g = [{"i": int(t["i"]), "x": float(t["x"]), "y": float(t["y"]), "z": float(t["z"])} for t in grabs]
        self._log("grabs", grabs=g)
        self.s.grabs = g

    def set(self, name, value):
        """A tuning field of the reference's Solver (gravity, viscosity, ...)."""
        self._log("set", name=name, value=value)
        setattr(self.s, name, value)

    def reset_solid(self):
        self._log("reset_solid")
        self.s.resetSolid()

    def clear_fluid(self):
        self._log("clear_fluid")
        self.s.clearFluid()

    def step(self, dt, substeps=5):
        self._log("step", dt=float(dt), substeps=int(substeps))
        self.s.step(dt, substeps)
        self.frame += 1

    def bundle(self):
        return {"version": 1, "type": BUNDLE_TYPE, "setup": self.setup,
                "state": base64.b64encode(self.state).decode("ascii"), "frames": self.frame,
                "actions": list(self.actions)}


def solver_for(bundle, device=-1):
    """A solver of the recording's shape, in its starting state."""
    if bundle.get("type") != BUNDLE_TYPE or bundle.get("version") != 1:
        raise ValueError("not a version 1 fluidphysics recording")
    st = bundle["setup"]
    s = Solver(st["max_fluid"], params=st["params"], device=device)
    s.addSolidCube(*st["cube"])
    s.load_state(base64.b64decode(bundle["state"]))
    return s


def replay(bundle, device=-1, on_frame=None):
    """Feed the recorded inputs into a solver restored from the recording's
    state; on_frame(frame, solver) after every step.  Returns the solver."""
    s = solver_for(bundle, device)
    frame = 0
    for a in bundle["actions"]:
        op = a["op"]
        if op == "emit":
            s.emit_batch(np.asarray(a["pos"]).reshape(-1, 3), np.asarray(a["vel"]).reshape(-1, 3),
                         np.asarray(a["mass"]))
        elif op == "cut":
            s.cut(*a["o"], *a["d1"], *a["d2"])
        elif op == "grabs":
            s.grabs = [dict(t) for t in a["grabs"]]
        elif op == "set":
            setattr(s, a["name"], a["value"])
        elif op == "reset_solid":
            s.resetSolid()
        elif op == "clear_fluid":
            s.clearFluid()
        elif op == "step":
            s.step(a["dt"], a["substeps"])
            frame += 1
            if on_frame is not None:
                on_frame(frame, s)
        else:
            raise ValueError("unknown recorded input %r" % op)
    return s
