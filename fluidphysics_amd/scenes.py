"""Scene builders for the single-scene workloads (the caller side of the solver).

The reference's caller is the browser app (src/main.js: the cube set-up at
:8-9, the hose pattern `sprayStep` at :224-251, the grab selection at
:288-305, the knife at :314-342).  These builders produce the same kind of
calls -- addSolidCube, a batch of emitFluid, slice wedges, grab targets, hose
bursts -- for the three single-scene workloads, deterministically from a seed
(a counter-based splitmix64 stream, so every host and language can regenerate
the identical inputs).

  plain_cpu_scene   7^3 = 343-particle cube + 2,600 fluid, reference constants
  single_gpu_scene  20^3 = 8,000 cube + 65,536 fluid column, GPU-scale constants
  large_scene       24^3 = 13,824 cube + 262,144 fluid slab + 120/frame hose
  batched_scene     one of many independent scenes: cube 6^3..20^3, 1k..64k
                    fluid, tilted gravity, seeded slice wedges

GPU-scale constants follow deriveOpts(dim, maxFluid, 30 mm)
(src/gpu/main.js:30-54) and the GPU defaults (src/gpu/sim.js:34-63), with the
CPU solver's semantics (multiplicative break strain, positional grab blend).
"""
import math
from dataclasses import dataclass, field

import numpy as np

from ._abi import REFERENCE_DEFAULTS

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, offset=0):
    """Values k = offset .. offset+count-1 of the splitmix64 sequence of `seed`."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, count, offset=0):
    """Doubles in [0, 1) from the top 53 bits."""
    return (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


@dataclass
class Wedge:
    """A knife stroke: solver.cut(o, d1, d2) (src/solver.js:458)."""
    frame: int
    o: tuple
    d1: tuple
    d2: tuple


@dataclass
class Scene:
    name: str
    params: dict
    max_fluid: int
    cube: tuple                      # (cx, cy, cz, n, spacing, mass)
    fluid_pos: np.ndarray            # (Nf0, 3) float64, emitted in order before frame 0
    fluid_vel: np.ndarray
    fluid_mass: float
    frames: int
    substeps: int = 5
    dt: float = 1.0 / 60.0
    wedges: list = field(default_factory=list)
    grab: dict = None                # {'frames': (f0, f1), 'count': G, 'point': xyz, 'amp': A, 'omega': w}
    hose: dict = None                # {'per_frame': k, 'origin', 'target', 'speed', 'cone_deg', 'seed'}

    @property
    def solid_count(self):
        return self.cube[3] ** 3

    @property
    def capacity(self):
        return self.solid_count + self.max_fluid

    def params_with(self, **kw):
        p = dict(REFERENCE_DEFAULTS)
        p.update(self.params)
        p.update(kw)
        return p


def _jittered_block(nx, ny, nz, spacing, origin, jitter, seed):
    """Lattice block (x-major, then y, then z) with uniform jitter +-jitter per axis."""
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    base = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64) * spacing
    base += np.asarray(origin, dtype=np.float64)
    u = uniform01(seed, base.size).reshape(-1, 3)
    return base + (u * 2.0 - 1.0) * jitter


def _normalize(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.sqrt(np.sum(v * v))


def plain_cpu_scene(frames=300, seed=1):
    """7^3 soft cube + 2,600 jittered fluid, the reference CPU page's scale.

    Cube at (0, 2, 0), n=7, spacing 0.18, mass 0.08 (src/main.js:8-9); fluid
    13 x 10 x 20 at spacing 0.18 from x=0.75 (0.19 clear of the cube's
    contact range), y=0.2, z centred, jitter +-0.018; one slice at frame 100
    from the camera (5.2, 3.6, 6.4) (src/main.js:28) between the rays to
    (0.1, 3.2, 0) and (0.1, 0, 0).  The wedge plane is kept off the cube's
    central lattice column: a plane through a lattice plane leaves particles
    exactly on it, where the side test (src/solver.js:472) flips on rounding.
    Jacobi relaxation 8 (3 iterations) reproduces the reference's
    Gauss-Seidel statistics with these constants.
    """
    sp = 0.18
    nx, ny, nz = 13, 10, 20
    origin = (0.75, 0.2, -(nz - 1) * sp * 0.5)
    fp = _jittered_block(nx, ny, nz, sp, origin, 0.1 * sp, seed)
    cam = np.array([5.2, 3.6, 6.4])
    d1 = _normalize(np.array([0.1, 3.2, 0.0]) - cam)
    d2 = _normalize(np.array([0.1, 0.0, 0.0]) - cam)
    return Scene(
        name="plain_cpu_343_2600",
        params=dict(gravity=-9.8, constraint_iters=3, jacobi_omega=8.0),
        max_fluid=2600,
        cube=(0.0, 2.0, 0.0, 7, sp, 0.08),
        fluid_pos=fp,
        fluid_vel=np.zeros_like(fp),
        fluid_mass=0.03,
        frames=frames,
        wedges=[Wedge(100, tuple(cam), tuple(d1), tuple(d2))],
    )


def gpu_scale_params(r=0.03):
    """deriveOpts(dim, maxFluid, 30 mm) constants (src/gpu/main.js:30-54) with the
    GPU defaults of src/gpu/sim.js:34-63, in the CPU solver's parameterisation."""
    H = 2.6 * r
    return dict(
        gravity=-9.8, solid_radius=r, fluid_radius=r, fluid_h=H, cell_size=H,
        stiffness=0.0015 * (H / 0.13), near_stiffness=0.006 * (H / 0.13),
        rest_density=3.0, compliance=5e-4, plastic_threshold=0.38, plastic_rate=0.04,
        viscosity=0.3, max_speed=14.0, solid_damping=1.2, fluid_damping=0.6,
        floor_friction_solid=0.4, floor_friction_fluid=0.04, cull_radius=7.0,
        grab_stiffness=1.0, constraint_iters=3, jacobi_omega=4.0,
    )


def _random_wedge(frame, seed, k, center, radius):
    """A stroke from a seeded camera position on a ring around `center`,
    sweeping across the body: rays to two points on either side of it."""
    u = uniform01(seed, 6, offset=16 * k)
    ang = 2.0 * math.pi * u[0]
    cam = np.array([center[0] + 4.0 * math.cos(ang), center[1] + 1.5 + u[1], center[2] + 4.0 * math.sin(ang)])
    side = _normalize(np.cross(np.asarray(center) - cam, [0.0, 1.0, 0.0]))
    off = (u[2] - 0.5) * radius
    a = np.asarray(center) + side * off + np.array([0.0, 1.5 * radius, 0.0])
    b = np.asarray(center) + side * (off + (u[3] - 0.5) * radius) - np.array([0.0, 1.5 * radius, 0.0])
    return Wedge(frame, tuple(cam), tuple(_normalize(a - cam)), tuple(_normalize(b - cam)))


def single_gpu_scene(frames=600, seed=2):
    """20^3 cube + 65,536 fluid column, two seeded wedges, a 64-particle grab."""
    r = 0.03
    sp = 1.91 * r
    dim = 20
    cy = (dim - 1) * sp / 2 + 0.5
    fs = 0.041  # near rest density 3 at H = 0.078, like the large slab
    nx, ny, nz = 32, 64, 32
    origin = (0.7, r + 0.01, -(nz - 1) * fs * 0.5)
    fp = _jittered_block(nx, ny, nz, fs, origin, 0.1 * fs, seed)
    center = (0.0, cy, 0.0)
    rad = (dim - 1) * sp * 0.5
    return Scene(
        name="single_gpu_8000_65536",
        params=gpu_scale_params(r),
        max_fluid=65536,
        cube=(0.0, cy, 0.0, dim, sp, 100 * r ** 3),
        fluid_pos=fp,
        fluid_vel=np.zeros_like(fp),
        fluid_mass=12 * r ** 3,
        frames=frames,
        wedges=[_random_wedge(150, seed, 0, center, rad), _random_wedge(350, seed, 1, center, rad)],
        grab=dict(frames=(400, 520), count=64, point=(0.0, cy + rad, 0.0), amp=0.25, omega=2.0 * math.pi * 0.5),
    )


def large_scene(frames=600, seed=3):
    """24^3 cube + 262,144 fluid slab (40% of the 7 x 7 x 0.92 basin) + hose."""
    r = 0.03
    sp = 1.91 * r
    dim = 24
    half = (dim - 1) * sp * 0.5
    fs = 0.041
    nx, ny, nz = 128, 16, 128
    origin = (-(nx - 1) * fs * 0.5, r + 0.01, -(nz - 1) * fs * 0.5)
    fp = _jittered_block(nx, ny, nz, fs, origin, 0.1 * fs, seed)
    top = origin[1] + (ny - 1) * fs
    cy = top + 0.35 + half
    per_frame = 120
    return Scene(
        name="large_13824_262144_hose",
        params=gpu_scale_params(r),
        max_fluid=nx * ny * nz + frames * per_frame,
        cube=(0.0, cy, 0.0, dim, sp, 100 * r ** 3),
        fluid_pos=fp,
        fluid_vel=np.zeros_like(fp),
        fluid_mass=12 * r ** 3,
        frames=frames,
        hose=dict(per_frame=per_frame, origin=(-2.9, 1.6, -2.9), target=(0.0, 0.4, 0.0), speed=4.0,
                  cone_deg=10.0, seed=seed + 100, radius=0.12),
    )


def batched_scene(index, seed=4, frames=600):
    """Scene `index` of the batched workload: cube side uniform in 6..20,
    fluid count log-uniform over [1000, 65536], gravity tilted by up to 15
    degrees towards a uniform azimuth, one or two seeded slice wedges.  Every
    draw comes from the splitmix64 stream of (seed, index), so a scene does
    not depend on how many others are generated."""
    r = 0.03
    sp = 1.91 * r
    fs = 0.041
    u = uniform01(seed * 1000003 + index, 8)
    dim = 6 + min(14, int(u[0] * 15))
    nf = int(round(math.exp(math.log(1000.0) + u[1] * (math.log(65536.0) - math.log(1000.0)))))
    tilt = math.radians(15.0) * u[2]
    azim = 2.0 * math.pi * u[3]
    g = 9.8
    ny = max(4, min(32, int(round(nf ** (1.0 / 3.0)))))
    nxz = int(math.ceil(math.sqrt(nf / ny)))
    origin = (0.7, r + 0.01, -(nxz - 1) * fs * 0.5)
    fp = _jittered_block(nxz, ny, nxz, fs, origin, 0.1 * fs, seed * 7919 + index)[:nf]
    cy = (dim - 1) * sp / 2 + 0.5
    center = (0.0, cy, 0.0)
    rad = (dim - 1) * sp * 0.5
    wseed = seed * 104729 + index
    wedges = [_random_wedge(20 + int(u[4] * min(200, frames)), wseed, 0, center, rad)]
    if u[5] < 0.5:
        wedges.append(_random_wedge(wedges[0].frame + 30 + int(u[6] * 200), wseed, 1, center, rad))
    params = gpu_scale_params(r)
    params.update(gravity=-g * math.cos(tilt), gravity_x=g * math.sin(tilt) * math.cos(azim),
                  gravity_z=g * math.sin(tilt) * math.sin(azim))
    return Scene(
        name="batched_%d_%d_%d" % (index, dim ** 3, nf),
        params=params,
        max_fluid=nf,
        cube=(0.0, cy, 0.0, dim, sp, 100 * r ** 3),
        fluid_pos=fp,
        fluid_vel=np.zeros_like(fp),
        fluid_mass=12 * r ** 3,
        frames=frames,
        wedges=[w for w in wedges if w.frame < frames],
    )


def batched_scenes(count=256, seed=4, frames=600):
    """The batched workload: `count` independent scenes (see batched_scene)."""
    return [batched_scene(k, seed, frames) for k in range(count)]


@dataclass
class Request:
    """One serving request: scene `index` of the batched workload (seed
    `seed`), stepped for `horizon` frames from its arrival."""
    id: int
    arrival: float       # seconds after the start of the stream (or a frame index)
    index: int
    seed: int
    horizon: int

    def scene(self):
        return batched_scene(self.index, self.seed, frames=self.horizon)


def serving_requests(rate_hz, duration_s, seed=5):
    """Open-loop Poisson stream: exponential inter-arrival gaps of mean
    1/rate_hz over duration_s seconds, each request a batched-workload scene
    (cube 6^3..20^3, 1k..64k fluid, tilted gravity, slices) with a horizon
    log-uniform over 1..600 frames."""
    out = []
    t = 0.0
    k = 0
    while True:
        u = uniform01(seed, 3, offset=3 * k)
        t += -math.log(1.0 - u[0]) / rate_hz
        if t >= duration_s:
            return out
        horizon = max(1, min(600, int(round(math.exp(u[1] * math.log(600.0))))))
        out.append(Request(id=k, arrival=t, index=int(u[2] * (1 << 30)), seed=seed, horizon=horizon))
        k += 1


def hose_burst(hose, frame, spacing):
    """Particles the hose emits at `frame`: positions on a golden-angle disc
    across the nozzle (src/main.js:235-241 pattern), directions uniform in a
    cone of hose['cone_deg'] around the axis, speed hose['speed'].  Returns
    (pos (k,3), vel (k,3)) float64."""
    k = hose["per_frame"]
    o = np.asarray(hose["origin"], dtype=np.float64)
    axis = _normalize(np.asarray(hose["target"], dtype=np.float64) - o)
    side = _normalize(np.cross(axis, [0.0, 1.0, 0.0]))
    up = _normalize(np.cross(side, axis))
    idx = np.arange(k, dtype=np.float64)
    phase = (frame * k + idx) * 2.39996
    rad = hose["radius"] * np.sqrt((idx + 0.5) / k)
    depth = (idx % 4) * spacing
    pos = o + side * (np.cos(phase) * rad)[:, None] + up * (np.sin(phase) * rad)[:, None] + axis * depth[:, None]
    u = uniform01(hose["seed"], 2 * k, offset=2 * k * frame).reshape(k, 2)
    cmax = math.cos(math.radians(hose["cone_deg"]))
    cth = 1.0 - u[:, 0] * (1.0 - cmax)
    sth = np.sqrt(np.maximum(0.0, 1.0 - cth * cth))
    phi = 2.0 * math.pi * u[:, 1]
    d = axis * cth[:, None] + side * (sth * np.cos(phi))[:, None] + up * (sth * np.sin(phi))[:, None]
    return pos, d * hose["speed"]


def grab_targets(scene, frame, ids, rest_pos):
    """Kinematic grab of `ids` on a sinusoidal path: targets = rest positions
    + amp * (sin wt, |sin wt|/2, cos wt - 1) with t counted from the grab start."""
    g = scene.grab
    t = (frame - g["frames"][0]) * scene.dt
    w = g["omega"]
    off = g["amp"] * np.array([math.sin(w * t), 0.5 * abs(math.sin(w * t)), math.cos(w * t) - 1.0])
    return rest_pos + off


SCENES = {
    "plain_cpu": plain_cpu_scene,
    "single_gpu": single_gpu_scene,
    "large": large_scene,
}
