"""ctypes wrapper of the oracle (TEST INFRASTRUCTURE ONLY).

The oracle is the plain-C restatement of the reference CPU solver
(src/solver.js) in oracle/fp_oracle.c.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may use it -- as the checker, never as the
thing measured or shipped.

    OracleSolver(params, max_fluid, storage="f32" | "f64", mode="gs" | "jacobi")

storage "f32" reproduces node's solver.js bit for bit (Float32Array stores),
"f64" is the FP64 replay.  mode "gs" is the reference's Gauss-Seidel order,
"jacobi" the formulation the HIP kernels implement.
"""
import ctypes
import os
import subprocess

import numpy as np

from fluidphysics_amd._abi import FpParams, make_params

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def build():
    """Compile the oracle libraries (make in oracle/)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib(storage):
    if storage in _LIBS:
        return _LIBS[storage]
    path = os.path.join(_HERE, "libfp_oracle%s.so" % {"f32": "32", "f64": "64"}[storage])
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
    FP = ctypes.POINTER(ctypes.c_float)
    DP = ctypes.POINTER(ctypes.c_double)
    IP = ctypes.POINTER(ctypes.c_int32)
    U8 = ctypes.POINTER(ctypes.c_uint8)
    sig = {
        "or_create": (P, [ctypes.POINTER(FpParams), I]),
        "or_destroy": (None, [P]),
        "or_set_params": (None, [P, ctypes.POINTER(FpParams)]),
        "or_set_mode": (None, [P, I]),
        "or_add_solid_cube": (I, [P, D, D, D, I, D, D]),
        "or_reset_solid": (None, [P]),
        "or_emit_batch": (None, [P, DP, DP, DP, I]),
        "or_clear_fluid": (None, [P]),
        "or_fluid_count": (I, [P]),
        "or_set_grabs": (None, [P, IP, DP, I]),
        "or_build_grid": (None, [P]),
        "or_substep": (None, [P, D]),
        "or_substep_until_contacts": (None, [P, D]),
        "or_solve_contacts": (None, [P]),
        "or_solve_constraints": (None, [P, D, I]),
        "or_collect_contacts": (None, [P]),
        "or_solid_extra_step": (None, [P]),
        "or_tet_update": (None, [P]),
        "or_num_tets": (I, [P]),
        "or_read_tets": (None, [P, IP, DP, U8]),
        "or_write_tet_alive": (None, [P, U8]),
        "or_read_cache": (None, [P, IP, IP, IP]),
        "or_step": (None, [P, D, I]),
        "or_recompute_groups": (None, [P]),
        "or_cut": (I, [P, D, D, D, D, D, D, D, D, D]),
        "or_pick_solid": (I, [P, D, D, D, D, D, D, D, DP]),
        "or_get_info": (None, [P, IP]),
        "or_read_state": (None, [P, FP, FP, FP, FP, U8, IP]),
        "or_read_state64": (None, [P, DP, DP, DP]),
        "or_write_state": (None, [P, FP, FP, FP, FP, U8]),
        "or_read_grid": (None, [P, IP, IP]),
        "or_read_constraints": (None, [P, IP, IP, DP, U8]),
        "or_write_constraints": (None, [P, DP, U8]),
        "or_read_density": (None, [P, FP, FP]),
        "or_read_pairs": (None, [P, IP, IP]),
        "or_hash_cell": (ctypes.c_uint32, [P, D, D, D]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIBS[storage] = lib
    return lib


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class OracleSolver:
    def __init__(self, params=None, max_fluid=2600, storage="f32", mode="gs"):
        self.lib = _lib(storage)
        self.storage = storage
        if isinstance(params, FpParams):
            self.params = params
        else:
            self.params = make_params(**(params or {}))
        self.h = self.lib.or_create(ctypes.byref(self.params), int(max_fluid))
        if not self.h:
            raise MemoryError("or_create failed")
        self.set_mode(mode)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.or_destroy(h)
            self.h = None

    def set_mode(self, mode):
        self.mode = mode
        self.lib.or_set_mode(self.h, {"gs": 0, "jacobi": 1}[mode])

    def set_params(self, **kw):
        self.params.update(**kw)
        self.lib.or_set_params(self.h, ctypes.byref(self.params))

    def add_solid_cube(self, cx, cy, cz, n=7, spacing=0.18, mass=0.08):
        if self.lib.or_add_solid_cube(self.h, cx, cy, cz, int(n), spacing, mass):
            raise RuntimeError("or_add_solid_cube failed")

    def info(self):
        buf = (ctypes.c_int32 * 8)()
        self.lib.or_get_info(self.h, buf)
        keys = ["capacity", "solid_count", "max_fluid", "table_size", "num_constraints",
                "fluid_next", "pair_count", "components_dirty"]
        return dict(zip(keys, list(buf)))

    def emit(self, pos, vel=None, mass=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        vel = np.zeros_like(pos) if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        if mass is None:
            m = None
        else:
            m = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, dtype=np.float64), (n,)))
        self.lib.or_emit_batch(self.h, _ptr(pos, ctypes.c_double), _ptr(vel, ctypes.c_double),
                               _ptr(m, ctypes.c_double), n)

    def reset_solid(self):
        self.lib.or_reset_solid(self.h)

    def clear_fluid(self):
        self.lib.or_clear_fluid(self.h)

    def fluid_count(self):
        return self.lib.or_fluid_count(self.h)

    def set_grabs(self, idx, xyz):
        idx = np.ascontiguousarray(idx, dtype=np.int32).ravel()
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
        self.lib.or_set_grabs(self.h, _ptr(idx, ctypes.c_int32), _ptr(xyz, ctypes.c_double), idx.size)

    def build_grid(self):
        self.lib.or_build_grid(self.h)

    def substep(self, h):
        self.lib.or_substep(self.h, h)

    def substep_until_contacts(self, h):
        self.lib.or_substep_until_contacts(self.h, h)

    def solve_contacts(self):
        """The contact pass alone (grid of the current positions, then
        contacts in the current mode)."""
        self.lib.or_solve_contacts(self.h)

    def solve_constraints(self, h, iters=1):
        """`iters` constraint sweeps (gs) or iterations (jacobi) alone, with
        the compliance of substep h."""
        self.lib.or_solve_constraints(self.h, h, int(iters))

    # ---- the GPU twin's optional solid modes (tets, contact cache) ------
    def collect_contacts(self):
        self.lib.or_collect_contacts(self.h)

    def solid_extra_step(self):
        self.lib.or_solid_extra_step(self.h)

    def tet_update(self):
        self.lib.or_tet_update(self.h)

    def read_tets(self):
        nt = self.lib.or_num_tets(self.h)
        v = np.zeros(4 * nt, np.int32)
        v0 = np.zeros(nt, np.float64)
        al = np.zeros(nt, np.uint8)
        self.lib.or_read_tets(self.h, _ptr(v, ctypes.c_int32), _ptr(v0, ctypes.c_double), _ptr(al, ctypes.c_uint8))
        return dict(v=v.reshape(-1, 4), v0=v0, alive=al)

    def write_tet_alive(self, alive):
        a = np.ascontiguousarray(alive, dtype=np.uint8)
        self.lib.or_write_tet_alive(self.h, _ptr(a, ctypes.c_uint8))

    def read_cache(self):
        ns = self.info()["solid_count"]
        ids = np.zeros(ns * 16, np.int32)
        n = np.zeros(ns, np.int32)
        ov = np.zeros(1, np.int32)
        self.lib.or_read_cache(self.h, _ptr(ids, ctypes.c_int32), _ptr(n, ctypes.c_int32), _ptr(ov, ctypes.c_int32))
        return dict(ids=ids.reshape(ns, 16), n=n, overflow=int(ov[0]))

    def step(self, dt=1.0 / 60.0, substeps=5):
        self.lib.or_step(self.h, dt, int(substeps))

    def recompute_groups(self):
        self.lib.or_recompute_groups(self.h)

    def cut(self, o, d1, d2):
        return self.lib.or_cut(self.h, *[float(x) for x in (*o, *d1, *d2)])

    def pick_solid(self, o, d, max_perp=0.2):
        t = ctypes.c_double()
        i = self.lib.or_pick_solid(self.h, *[float(x) for x in (*o, *d)], float(max_perp), ctypes.byref(t))
        return None if i < 0 else (i, t.value)

    def hash_cell(self, x, y, z):
        return self.lib.or_hash_cell(self.h, float(x), float(y), float(z))

    def read_state(self):
        inf = self.info()
        n, ns = inf["capacity"], inf["solid_count"]
        pos = np.zeros(3 * n, np.float32)
        prev = np.zeros(3 * n, np.float32)
        vel = np.zeros(3 * n, np.float32)
        im = np.zeros(n, np.float32)
        act = np.zeros(n, np.uint8)
        grp = np.zeros(max(ns, 1), np.int32)
        self.lib.or_read_state(self.h, _ptr(pos, ctypes.c_float), _ptr(prev, ctypes.c_float),
                               _ptr(vel, ctypes.c_float), _ptr(im, ctypes.c_float),
                               _ptr(act, ctypes.c_uint8), _ptr(grp, ctypes.c_int32))
        return dict(pos=pos, prev=prev, vel=vel, inv_mass=im, active=act, group=grp[:ns])

    def read_state64(self):
        n = self.info()["capacity"]
        pos = np.zeros(3 * n, np.float64)
        prev = np.zeros(3 * n, np.float64)
        vel = np.zeros(3 * n, np.float64)
        self.lib.or_read_state64(self.h, _ptr(pos, ctypes.c_double), _ptr(prev, ctypes.c_double),
                                 _ptr(vel, ctypes.c_double))
        return dict(pos=pos, prev=prev, vel=vel)

    def write_state(self, pos=None, prev=None, vel=None, inv_mass=None, active=None):
        def f(a, dt):
            return None if a is None else np.ascontiguousarray(a, dtype=dt).ravel()
        pos, prev, vel, inv_mass = (f(a, np.float32) for a in (pos, prev, vel, inv_mass))
        active = f(active, np.uint8)
        self.lib.or_write_state(self.h, _ptr(pos, ctypes.c_float), _ptr(prev, ctypes.c_float),
                                _ptr(vel, ctypes.c_float), _ptr(inv_mass, ctypes.c_float),
                                _ptr(active, ctypes.c_uint8))

    def read_grid(self):
        inf = self.info()
        cs = np.zeros(inf["table_size"] + 1, np.int32)
        ce = np.zeros(inf["capacity"], np.int32)
        self.lib.or_read_grid(self.h, _ptr(cs, ctypes.c_int32), _ptr(ce, ctypes.c_int32))
        return cs, ce

    def read_constraints(self):
        nc = self.info()["num_constraints"]
        a = np.zeros(nc, np.int32)
        b = np.zeros(nc, np.int32)
        rest = np.zeros(nc, np.float64)
        alive = np.zeros(nc, np.uint8)
        self.lib.or_read_constraints(self.h, _ptr(a, ctypes.c_int32), _ptr(b, ctypes.c_int32),
                                     _ptr(rest, ctypes.c_double), _ptr(alive, ctypes.c_uint8))
        return dict(a=a, b=b, rest=rest, alive=alive)

    def write_constraints(self, rest=None, alive=None):
        rest = None if rest is None else np.ascontiguousarray(rest, dtype=np.float64)
        alive = None if alive is None else np.ascontiguousarray(alive, dtype=np.uint8)
        self.lib.or_write_constraints(self.h, _ptr(rest, ctypes.c_double), _ptr(alive, ctypes.c_uint8))

    def read_density(self):
        n = self.info()["capacity"]
        rho = np.zeros(n, np.float32)
        rn = np.zeros(n, np.float32)
        self.lib.or_read_density(self.h, _ptr(rho, ctypes.c_float), _ptr(rn, ctypes.c_float))
        return rho, rn

    def read_pairs(self):
        npair = self.info()["pair_count"]
        pi = np.zeros(npair, np.int32)
        pj = np.zeros(npair, np.int32)
        self.lib.or_read_pairs(self.h, _ptr(pi, ctypes.c_int32), _ptr(pj, ctypes.c_int32))
        return pi, pj


def from_scene(scene, storage="f32", mode="gs", **param_overrides):
    """An oracle solver set up with a scene's cube and initial fluid."""
    s = OracleSolver(scene.params_with(**param_overrides), scene.max_fluid, storage=storage, mode=mode)
    s.add_solid_cube(*scene.cube)
    if len(scene.fluid_pos):
        s.emit(scene.fluid_pos, scene.fluid_vel, scene.fluid_mass)
    return s
