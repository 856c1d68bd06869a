"""GPU-side helpers for the parity tests: the product solver (libfluidphysics.so
through fluidphysics_amd.Solver) behind the same small surface the oracle
wrapper offers, so one driver runs both."""
import sys

import numpy as np

from fluidphysics_amd import Solver


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class GpuAdapter:
    def __init__(self, solver):
        self.s = solver

    def cut(self, o, d1, d2):
        return self.s.cut(*o, *d1, *d2)

    def pick_solid(self, o, d, max_perp=0.2):
        hit = self.s.pickSolid(*o, *d, max_perp)
        return None if hit is None else (hit["id"], hit["t"])

    def set_grabs(self, idx, xyz):
        xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
        self.s.grabs = [{"i": int(i), "x": float(p[0]), "y": float(p[1]), "z": float(p[2])} for i, p in zip(idx, xyz)]

    def step(self, dt, substeps):
        self.s.step(dt, substeps)

    def read_state(self):
        return self.s.read_state()


def make_gpu(d, iters=3, omega=4.0, **overrides):
    """A GPU solver for a golden scenario dict (tests/golden/gen_inputs.py);
    iters / omega None keep the library defaults."""
    params = dict(gravity=d["gravity"])
    if iters is not None:
        params["constraint_iters"] = iters
    if omega is not None:
        params["jacobi_omega"] = omega
    params.update(overrides)
    s = Solver(d["maxFluid"], **params)
    s.addSolidCube(*d["cube"])
    e = np.asarray(d["emit"], dtype=np.float64)
    if len(e):
        s.emit_batch(e[:, 0:3], e[:, 3:6], e[:, 6])
    return s


def tol_report(name, a, b, scale):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    err = np.abs(a - b)
    m = float(err.max()) if err.size else 0.0
    log("  %-28s max|diff| = %.3e  (tolerance %.3e)" % (name, m, scale))
    return m
