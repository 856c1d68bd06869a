"""MI355X-native unified-particle solver: host mirror of the reference's
`class Solver` (src/solver.js:8-516) over the C ABI of libfluidphysics.so.

    from fluidphysics_amd import Solver, FLUID, SOLID
    s = Solver(maxFluid=2600)
    s.addSolidCube(0, 2.0, 0, 7, 0.18)
    s.emitFluid(x, y, z, vx, vy, vz)          # or s.emit_batch(pos, vel, mass)
    s.grabs = [{"i": 12, "x": 0.1, "y": 2.0, "z": 0.0}]
    s.step(1 / 60, 5)
    n = s.cut(ox, oy, oz, d1x, d1y, d1z, d2x, d2y, d2z)
    hit = s.pickSolid(ox, oy, oz, dx, dy, dz, 0.2)   # {"id", "t"} or None
    s.pos, s.vel, s.active, s.group                 # numpy views, refreshed lazily

Method names, defaults, return shapes and the mutable tuning fields are the
reference's.  Differences: errors from the device raise FpError (the
reference has none to raise); emitFluid calls are buffered and sent as one
batch before the next operation that needs them.  All solving happens in the
HIP kernels of libfluidphysics.so; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from . import _abi
from ._abi import FpInfo, FpParams, FpPassStat, make_params

FLUID = 0
SOLID = 1


class FpError(RuntimeError):
    pass


# reference field name (src/solver.js:14-37) -> fp_params field
_FIELDS = {
    "gravity": "gravity",
    "solidRadius": "solid_radius",
    "fluidRadius": "fluid_radius",
    "compliance": "compliance",
    "plasticThreshold": "plastic_threshold",
    "plasticRate": "plastic_rate",
    "breakStrain": "break_strain",
    "fluidH": "fluid_h",
    "restDensity": "rest_density",
    "stiffness": "stiffness",
    "nearStiffness": "near_stiffness",
    "viscosity": "viscosity",
    "solidDamping": "solid_damping",
    "fluidDamping": "fluid_damping",
    "floorFrictionSolid": "floor_friction_solid",
    "floorFrictionFluid": "floor_friction_fluid",
    "cullRadius": "cull_radius",
    "maxSpeed": "max_speed",
    "grabStiffness": "grab_stiffness",
    "cellSize": "cell_size",
    "constraintIters": "constraint_iters",
    "gravityX": "gravity_x",
    "gravityZ": "gravity_z",
    "tetVolume": "tet_volume",
    "solidContactCache": "solid_contact_cache",
    "tearGap": "tear_gap",
}


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _ip(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _u8p(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


class Solver:
    """new Solver({ maxFluid }) (src/solver.js:9)."""

    def __init__(self, maxFluid=2600, params=None, device=-1, **param_overrides):
        lib = _abi.load_library()
        object.__setattr__(self, "_lib", lib)
        if isinstance(params, FpParams):
            p = params
        else:
            p = FpParams()
            lib.fp_params_default(ctypes.byref(p))
            if params:
                p.update(**params)
        p.update(**param_overrides)
        h = ctypes.c_void_p()
        self._check(lib.fp_create(ctypes.byref(p), int(maxFluid), int(device), ctypes.byref(h)), None)
        object.__setattr__(self, "_h", h)
        object.__setattr__(self, "_params", p)
        object.__setattr__(self, "maxFluid", int(maxFluid))
        object.__setattr__(self, "solidCount", 0)
        object.__setattr__(self, "capacity", 0)
        object.__setattr__(self, "grabs", [])
        object.__setattr__(self, "_pending_emit", [])
        object.__setattr__(self, "_cache", {})
        object.__setattr__(self, "_params_dirty", False)
        object.__setattr__(self, "_grabs_pushed", False)

    # ---- plumbing -------------------------------------------------------
    def _check(self, rc, h="self"):
        if rc != 0:
            lib = self._lib
            handle = None if h is None else self._h
            msg = lib.fp_last_error(handle)
            raise FpError("fluidphysics error %d: %s" % (rc, msg.decode() if msg else ""))

    def __del__(self):
        h = self.__dict__.get("_h")
        if h:
            self._lib.fp_destroy(h)
            self.__dict__["_h"] = None

    def close(self):
        self.__del__()

    def __getattr__(self, name):
        if name in _FIELDS:
            return getattr(self._params, _FIELDS[name])
        if name == "wall":
            p = self._params
            return {"x": p.wall_x, "z": p.wall_z, "height": p.wall_height, "halfThick": p.wall_half_thick}
        if name == "tableSize":
            return self.info().table_size
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in _FIELDS:
            setattr(self._params, _FIELDS[name], value)
            self.__dict__["_params_dirty"] = True
        elif name == "wall":
            p = self._params
            p.wall_x = value.get("x", p.wall_x)
            p.wall_z = value.get("z", p.wall_z)
            p.wall_height = value.get("height", p.wall_height)
            p.wall_half_thick = value.get("halfThick", p.wall_half_thick)
            self.__dict__["_params_dirty"] = True
        else:
            object.__setattr__(self, name, value)

    def _flush(self):
        """Send buffered tuning changes, emits and the grab list."""
        if self._params_dirty:
            self._check(self._lib.fp_set_params(self._h, ctypes.byref(self._params)))
            self.__dict__["_params_dirty"] = False
        if self._pending_emit:
            e = np.asarray(self._pending_emit, dtype=np.float64)
            self._pending_emit.clear()
            self.emit_batch(e[:, 0:3], e[:, 3:6], e[:, 6])

    def _push_grabs(self):
        g = self.grabs
        if not g and not self._grabs_pushed:
            return  # nothing held now or at the last step
        self.__dict__["_grabs_pushed"] = bool(g)
        idx = np.array([int(t["i"]) for t in g], np.int32)
        xyz = np.array([[float(t["x"]), float(t["y"]), float(t["z"])] for t in g], np.float64).reshape(-1, 3)
        xyz = np.ascontiguousarray(xyz)
        self._check(self._lib.fp_set_grabs(self._h, _ip(idx), _dp(xyz) if len(g) else None, idx.size))

    def _invalidate(self):
        self._cache.clear()

    def params(self):
        return self._params.to_dict()

    def info(self):
        inf = FpInfo()
        self._check(self._lib.fp_get_info(self._h, ctypes.byref(inf)))
        return inf

    # ---- the reference surface -------------------------------------------
    def addSolidCube(self, cx, cy, cz, n=7, spacing=0.18, particleMass=0.08):
        """src/solver.js:72-123."""
        self._flush()
        self._check(self._lib.fp_add_solid_cube(self._h, float(cx), float(cy), float(cz), int(n),
                                                float(spacing), float(particleMass)))
        inf = self.info()
        object.__setattr__(self, "solidCount", inf.solid_count)
        object.__setattr__(self, "capacity", inf.capacity)
        self._invalidate()

    def emitFluid(self, x, y, z, vx, vy, vz, mass=0.03):
        """src/solver.js:139-149 (buffered; flushed before the next step/read)."""
        self._pending_emit.append((x, y, z, vx, vy, vz, mass))
        self._invalidate()

    def emit_batch(self, pos, vel=None, mass=0.03):
        """emitFluid over n particles in order, one device transfer."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        vel = np.zeros_like(pos) if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, dtype=np.float64), (n,)))
        self._check(self._lib.fp_emit(self._h, _dp(pos), _dp(vel), _dp(m), n))
        self._invalidate()

    def resetSolid(self):
        """src/solver.js:125-137."""
        self._flush()
        self._check(self._lib.fp_reset_solid(self._h))
        self.grabs = []
        self._invalidate()

    def clearFluid(self):
        """src/solver.js:151-153."""
        self._flush()
        self._check(self._lib.fp_clear_fluid(self._h))
        self._invalidate()

    def fluidCount(self):
        """src/solver.js:155-159."""
        self._flush()
        c = ctypes.c_int32()
        self._check(self._lib.fp_fluid_count(self._h, ctypes.byref(c)))
        return c.value

    def nonfinite(self):
        """True once a finished substep ended a particle with a non-finite
        position or velocity (fp_nonfinite; up to date after a sync, sticky
        until restore() or import_state())."""
        c = ctypes.c_int32()
        self._check(self._lib.fp_nonfinite(self._h, ctypes.byref(c)))
        return bool(c.value)

    def step(self, dt, substeps=5, wait=True):
        """src/solver.js:187-242.  wait=False leaves the final wait to sync()."""
        self._flush()
        self._push_grabs()
        fn = self._lib.fp_step if wait else self._lib.fp_step_async
        self._check(fn(self._h, float(dt), int(substeps)))
        self._invalidate()

    def sync(self):
        self._check(self._lib.fp_sync(self._h))

    def cut(self, ox, oy, oz, d1x, d1y, d1z, d2x, d2y, d2z):
        """src/solver.js:458-486; returns the number of constraints cut."""
        self._flush()
        o = np.array([ox, oy, oz], np.float64)
        d1 = np.array([d1x, d1y, d1z], np.float64)
        d2 = np.array([d2x, d2y, d2z], np.float64)
        n = ctypes.c_int32()
        self._check(self._lib.fp_cut(self._h, _dp(o), _dp(d1), _dp(d2), ctypes.byref(n)))
        self._invalidate()
        return n.value

    def pickSolid(self, ox, oy, oz, dx, dy, dz, maxPerp=0.2):
        """src/solver.js:502-515; {'id', 't'} or None."""
        self._flush()
        o = np.array([ox, oy, oz], np.float64)
        d = np.array([dx, dy, dz], np.float64)
        i = ctypes.c_int32()
        t = ctypes.c_double()
        self._check(self._lib.fp_pick_solid(self._h, _dp(o), _dp(d), float(maxPerp), ctypes.byref(i), ctypes.byref(t)))
        return None if i.value < 0 else {"id": i.value, "t": t.value}

    def select_solids(self, center, radius=0.35, same_group_as=-1):
        """The grab selection of the reference's caller (src/main.js:293-301):
        ids of the solids in solid same_group_as's component (any when -1)
        within radius of center, ascending (device pass fp_select_solids)."""
        self._flush()
        c = np.asarray(center, np.float64).reshape(3)
        n = ctypes.c_int32()
        self._check(self._lib.fp_select_solids(self._h, _dp(c), float(radius), int(same_group_as), None, 0,
                                               ctypes.byref(n)))
        ids = np.zeros(max(n.value, 1), np.int32)
        self._check(self._lib.fp_select_solids(self._h, _dp(c), float(radius), int(same_group_as), _ip(ids), n.value,
                                               ctypes.byref(n)))
        return ids[: n.value]

    def components(self):
        """_recomputeGroups (src/solver.js:488-499) on the device over the
        constraints alive now (fp_components): (labels, count), labels[i] the
        smallest solid index of solid i's component.  The partition is the
        reference's; `group` keeps the reference's own root labels."""
        self._flush()
        labels = np.zeros(max(self.solidCount, 1), np.int32)
        n = ctypes.c_int32()
        self._check(self._lib.fp_components(self._h, _ip(labels), ctypes.byref(n)))
        return labels[: self.solidCount], n.value

    # ---- state arrays (src/solver.js:42-48), fetched on access -----------
    def read_state(self):
        self._flush()
        n, ns = self.capacity, self.solidCount
        out = dict(pos=np.zeros(3 * n, np.float32), prev=np.zeros(3 * n, np.float32),
                   vel=np.zeros(3 * n, np.float32), inv_mass=np.zeros(n, np.float32),
                   active=np.zeros(n, np.uint8), group=np.zeros(max(ns, 1), np.int32))
        self._check(self._lib.fp_read_state(self._h, _fp(out["pos"]), _fp(out["prev"]), _fp(out["vel"]),
                                            _fp(out["inv_mass"]), _u8p(out["active"]), _ip(out["group"])))
        out["group"] = out["group"][:ns]
        self._cache.update(out)
        return out

    # ---- checkpoints and state blobs (src/gpu/sim.js:494-547) ------------
    def checkpoint(self, slot=0):
        """Device-resident copy of the scene's state (slots 0..7)."""
        self._flush()
        self._check(self._lib.fp_checkpoint_save(self._h, int(slot)))

    def restore(self, slot=0):
        self._flush()
        self._check(self._lib.fp_checkpoint_restore(self._h, int(slot)))
        self._invalidate()

    def save_state(self):
        """The state as a portable blob (bytes): header + device arrays."""
        self._flush()
        n = ctypes.c_int64()
        self._check(self._lib.fp_state_size(self._h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        self._check(self._lib.fp_state_export(self._h, buf, n.value))
        return buf.raw

    def load_state(self, blob):
        self._flush()
        b = bytes(blob)
        self._check(self._lib.fp_state_import(self._h, b, len(b)))
        self._invalidate()

    def read_positions(self):
        """pos alone (fp_read_state with the other outputs omitted), (capacity, 3)."""
        self._flush()
        pos = np.zeros(3 * self.capacity, np.float32)
        self._check(self._lib.fp_read_state(self._h, _fp(pos), None, None, None, None, None))
        return pos.reshape(-1, 3)

    def _arr(self, key):
        if key not in self._cache:
            self.read_state()
        return self._cache[key]

    @property
    def pos(self):
        return self._arr("pos")

    @property
    def prev(self):
        return self._arr("prev")

    @property
    def vel(self):
        return self._arr("vel")

    @property
    def invMass(self):
        return self._arr("inv_mass")

    @property
    def active(self):
        return self._arr("active")

    @property
    def group(self):
        return self._arr("group")

    @property
    def type(self):
        t = np.zeros(self.capacity, np.uint8)
        t[: self.solidCount] = SOLID
        return t

    def write_state(self, pos=None, prev=None, vel=None, inv_mass=None, active=None):
        self._flush()

        def c(a, dt):
            return None if a is None else np.ascontiguousarray(a, dtype=dt).ravel()
        pos, prev, vel, inv_mass = (c(a, np.float32) for a in (pos, prev, vel, inv_mass))
        active = c(active, np.uint8)
        self._check(self._lib.fp_write_state(self._h, _fp(pos), _fp(prev), _fp(vel), _fp(inv_mass), _u8p(active)))
        self._invalidate()

    def read_constraints(self):
        nc = self.info().num_constraints
        a = np.zeros(nc, np.int32)
        b = np.zeros(nc, np.int32)
        rest = np.zeros(nc, np.float32)
        alive = np.zeros(nc, np.uint8)
        self._check(self._lib.fp_read_constraints(self._h, _ip(a), _ip(b), _fp(rest), _u8p(alive)))
        return dict(a=a, b=b, rest=rest, alive=alive)

    def write_constraints(self, rest=None, alive=None):
        rest = None if rest is None else np.ascontiguousarray(rest, dtype=np.float32)
        alive = None if alive is None else np.ascontiguousarray(alive, dtype=np.uint8)
        self._check(self._lib.fp_write_constraints(self._h, _fp(rest), _u8p(alive)))

    @property
    def consAlive(self):
        return self.read_constraints()["alive"]

    @property
    def consRest(self):
        return self.read_constraints()["rest"]

    def build_grid(self):
        """_buildGrid (src/solver.js:174-184) on the current positions."""
        self._flush()
        self._check(self._lib.fp_build_grid(self._h))

    def read_grid(self):
        inf = self.info()
        cs = np.zeros(inf.table_size + 1, np.int32)
        ce = np.zeros(inf.capacity, np.int32)
        self._check(self._lib.fp_read_grid(self._h, _ip(cs), _ip(ce)))
        return cs, ce

    def read_tets(self):
        """The optional tet mode's tets: (verts [NT, 4] int32, alive [NT] uint8)."""
        nt = ctypes.c_int32(0)
        self._check(self._lib.fp_read_tets(self._h, None, None, ctypes.byref(nt)))
        v = np.zeros((nt.value, 4), np.int32)
        a = np.zeros(nt.value, np.uint8)
        if nt.value:
            self._check(self._lib.fp_read_tets(self._h, _ip(v), _u8p(a), ctypes.byref(nt)))
        return v, a

    def read_cells(self, which=0):
        """The last substep's cell sort (0: grid order, 1: fluid order):
        dict(inv_cell, M, mx, my, slots, start [M+1], ids [slots], pos [slots, 4])."""
        desc = np.zeros(5, np.float64)
        self._check(self._lib.fp_read_cells(self._h, int(which), _dp(desc), None, None, None))
        M = int(desc[1])
        start = np.zeros(M + 1, np.int32)
        ids = np.zeros(self.capacity, np.int32)
        pos = np.zeros((self.capacity, 4), np.float32)
        self._check(self._lib.fp_read_cells(self._h, int(which), _dp(desc), _ip(start), _ip(ids), _fp(pos)))
        k = int(desc[4])
        return dict(inv_cell=float(desc[0]), M=M, mx=int(desc[2]), my=int(desc[3]), slots=k, start=start,
                    ids=ids[:k], pos=pos[:k])

    def read_density(self):
        n = self.capacity
        rho = np.zeros(n, np.float32)
        rn = np.zeros(n, np.float32)
        self._check(self._lib.fp_read_density(self._h, _fp(rho), _fp(rn)))
        return rho, rn

    # ---- measurement ------------------------------------------------------
    def profile(self, enable=True):
        self._check(self._lib.fp_profile_enable(self._h, 1 if enable else 0))

    def profile_read(self):
        buf = (FpPassStat * _abi.FP_MAX_PASSES)()
        n = ctypes.c_int32()
        self._check(self._lib.fp_profile_read(self._h, buf, _abi.FP_MAX_PASSES, ctypes.byref(n)))
        return {buf[k].name.decode(): (buf[k].launches, buf[k].total_ms) for k in range(n.value)}


class Batch:
    """Many independent scenes stepped together (fp_batch_*): one kernel
    launch per pass over every scene, a frame captured once into a HIP graph
    and replayed.  Scenes keep their own Solver API between frames; all must
    use the same constraintIters."""

    def __init__(self, device=-1):
        self._lib = _abi.load_library()
        h = ctypes.c_void_p()
        self._h = None
        self._check(self._lib.fp_batch_create(int(device), ctypes.byref(h)))
        self._h = h
        self.scenes = []

    def _check(self, rc):
        if rc != 0:
            msg = self._lib.fp_batch_last_error(self._h)
            raise FpError("fluidphysics batch error %d: %s" % (rc, msg.decode() if msg else ""))

    def add(self, solver):
        solver._flush()
        self._check(self._lib.fp_batch_add(self._h, solver._h))
        self.scenes.append(solver)

    def remove(self, solver):
        self._check(self._lib.fp_batch_remove(self._h, solver._h))
        self.scenes.remove(solver)

    def step(self, dt, substeps=5, wait=True):
        """One frame of every scene (src/solver.js:187-242 each)."""
        for s in self.scenes:
            s._flush()
            s._push_grabs()
        fn = self._lib.fp_batch_step if wait else self._lib.fp_batch_step_async
        self._check(fn(self._h, float(dt), int(substeps)))
        for s in self.scenes:
            s._invalidate()

    def sync(self):
        self._check(self._lib.fp_batch_sync(self._h))

    def graph(self, enable=True):
        self._check(self._lib.fp_batch_set_graph(self._h, 1 if enable else 0))

    def contact_lanes(self, lanes=0):
        """Pin the solid contact pass's lanes per solid (4..32; 0 = tune)."""
        self._check(self._lib.fp_batch_set_contact_lanes(self._h, int(lanes)))

    def info(self):
        inf = _abi.FpBatchInfo()
        self._check(self._lib.fp_batch_get_info(self._h, ctypes.byref(inf)))
        return inf

    def profile(self, enable=True):
        self._check(self._lib.fp_batch_profile_enable(self._h, 1 if enable else 0))

    def profile_read(self):
        buf = (FpPassStat * _abi.FP_MAX_PASSES)()
        n = ctypes.c_int32()
        self._check(self._lib.fp_batch_profile_read(self._h, buf, _abi.FP_MAX_PASSES, ctypes.byref(n)))
        return {buf[k].name.decode(): (buf[k].launches, buf[k].total_ms) for k in range(n.value)}

    def snapshot(self):
        """Positions of every scene after the frames queued so far
        (fp_batch_snapshot): a float32 view of 3 x capacity floats per scene,
        in pinned memory the batch owns and reuses two calls later -- copy
        what must outlive that -- and the scenes' offsets into it."""
        pos = ctypes.POINTER(ctypes.c_float)()
        off = ctypes.POINTER(ctypes.c_int64)()
        n = ctypes.c_int32()
        self._check(self._lib.fp_batch_snapshot(self._h, ctypes.byref(pos), ctypes.byref(off), ctypes.byref(n)))
        offsets = np.ctypeslib.as_array(off, shape=(n.value + 1,)).copy()
        total = int(offsets[-1])
        flat = np.ctypeslib.as_array(pos, shape=(total,)) if total else np.zeros(0, np.float32)
        return flat, offsets

    def positions(self):
        """snapshot() split per scene: one (capacity, 3) view per member."""
        flat, off = self.snapshot()
        return [flat[off[k]:off[k + 1]].reshape(-1, 3) for k in range(len(off) - 1)]

    def close(self):
        h = self.__dict__.get("_h")
        if h:
            self._lib.fp_batch_destroy(h)
            self._h = None
        self.scenes = []

    def __del__(self):
        self.close()


def scene_bytes(params, max_fluid, cube_n):
    """Device bytes a solver of these parameters holds after addSolidCube(n)
    and emits of up to max_fluid particles at once (fp_scene_bytes)."""
    lib = _abi.load_library()
    p = params if isinstance(params, FpParams) else make_params(**params)
    out = ctypes.c_int64()
    if lib.fp_scene_bytes(ctypes.byref(p), int(max_fluid), int(cube_n), ctypes.byref(out)) != 0:
        raise FpError("fp_scene_bytes: invalid arguments")
    return out.value


def pool_info(device=-1):
    """(reserved, cached) bytes of the device's memory pool."""
    lib = _abi.load_library()
    r, c = ctypes.c_int64(), ctypes.c_int64()
    lib.fp_pool_info(int(device), ctypes.byref(r), ctypes.byref(c))
    return r.value, c.value


def pool_trim(device=-1):
    _abi.load_library().fp_pool_trim(int(device))


def _last_error():
    lib = _abi.load_library()
    return lib.fp_last_error(None).decode(errors="replace")


def tune_cache_save(path):
    """Write the batches' tuning choices of this process (contact lanes and
    launch shape per group signature) to a JSON file (fp_tune_cache_save)."""
    if _abi.load_library().fp_tune_cache_save(os.fsencode(path)) != 0:
        raise FpError(_last_error())


def tune_cache_load(path):
    """Merge a tuning cache file (fp_tune_cache_load): later batches of a
    cached signature skip their tuning frames."""
    if _abi.load_library().fp_tune_cache_load(os.fsencode(path)) != 0:
        raise FpError(_last_error())


def tune_cache_reset():
    """Forget every tuning choice and stop using loaded contact lanes."""
    _abi.load_library().fp_tune_cache_reset()


def from_scene(scene, device=-1, **param_overrides):
    """A solver set up with a scene's cube and initial fluid (fluidphysics_amd.scenes)."""
    s = Solver(scene.max_fluid, params=scene.params_with(**param_overrides), device=device)
    s.addSolidCube(*scene.cube)
    if len(scene.fluid_pos):
        s.emit_batch(scene.fluid_pos, scene.fluid_vel, scene.fluid_mass)
    return s


__all__ = ["Solver", "Batch", "FLUID", "SOLID", "FpError", "FpParams", "make_params", "from_scene",
           "scene_bytes", "pool_info", "pool_trim", "tune_cache_save", "tune_cache_load", "tune_cache_reset"]
