"""Write the scenario inputs that tests/golden/make_golden.mjs feeds to the
reference solver (src/solver.js) to produce the golden vectors.

Run from the repo root:  python3 tests/golden/gen_inputs.py
(make_golden.sh runs both steps.)  The inputs are plain JSON numbers
(IEEE doubles, shortest round-trip repr), so node reads exactly the values
Python generated.
"""
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))

from fluidphysics_amd import scenes  # noqa: E402


def plain_cpu():
    sc = scenes.plain_cpu_scene()
    w = sc.wedges[0]
    return dict(
        name="plain_cpu",
        maxFluid=sc.max_fluid,
        gravity=-9.8,
        cube=list(sc.cube),
        emit=[[*map(float, p), 0.0, 0.0, 0.0, sc.fluid_mass] for p in sc.fluid_pos],
        frames=300,
        substeps=5,
        cuts=[dict(frame=w.frame, o=list(w.o), d1=list(w.d1), d2=list(w.d2))],
        dumpFrames=[150],
    )


def grab_tear():
    """Settle the plain scene 60 frames, pick the cube from the camera, grab the
    pick's group within 0.35 (src/main.js:288-305) and drag the targets away
    fast enough to tear constraints (src/solver.js:370)."""
    d = plain_cpu()
    d.update(name="grab_tear", frames=100, cuts=[], dumpFrames=[], dumpInit=False)
    cam = [5.2, 3.6, 6.4]
    tgt = [0.0, 0.6, 0.0]
    v = [tgt[k] - cam[k] for k in range(3)]
    n = math.sqrt(sum(x * x for x in v))
    d["grab"] = dict(frame=60, o=cam, d=[x / n for x in v], maxPerp=0.25, radius=0.35,
                     drag=[0.0, 1.5, 0.4], until=75)
    return d


def ring_cap():
    """Small ring (maxFluid 256) emitted 300 times: the ring wraps and
    overwrites (src/solver.js:140-141); a tight cluster overflows the pair cap
    32*maxFluid (:57, :311); fast particles leave the basin and are culled
    (:421-425); one particle sits at x = 1e7 where the hash products exceed
    2^53 (:163)."""
    rng_u = scenes.uniform01(7, 300 * 3).reshape(300, 3)
    emit = []
    for k in range(300):
        if k < 200:
            p = [0.6 + 0.03 * (k % 6) + 0.005 * rng_u[k, 0], 0.4 + 0.03 * ((k // 6) % 6),
                 -0.3 + 0.03 * (k // 36) + 0.005 * rng_u[k, 2]]
            v = [0.0, 0.0, 0.0]
        elif k < 290:
            p = [-1.0 + 0.2 * rng_u[k, 0], 1.0 + 0.2 * rng_u[k, 1], -1.0 + 0.2 * rng_u[k, 2]]
            v = [-12.0 * rng_u[k, 0] - 2.0, 3.0, -11.0 * rng_u[k, 2] - 2.0]
        else:
            p = [1e7 if k == 290 else 2.0 + 0.1 * (k - 290), 0.5, 0.5]
            v = [0.0, 0.0, 0.0]
        emit.append([*p, *v, 0.03])
    return dict(name="ring_cap", maxFluid=256, gravity=-10.0, cube=[0.0, 1.2, 0.0, 3, 0.18, 0.08],
                emit=emit, frames=30, substeps=5, cuts=[], dumpFrames=[30])


# the solver's public fields (src/solver.js:14-37, cellSize :50) for the
# GPU-scale constant set; fluidDamping stays the reference's hard-coded 0.6
_JS_FIELDS = dict(solid_radius="solidRadius", fluid_radius="fluidRadius", compliance="compliance",
                  plastic_threshold="plasticThreshold", plastic_rate="plasticRate", fluid_h="fluidH",
                  rest_density="restDensity", stiffness="stiffness", near_stiffness="nearStiffness",
                  viscosity="viscosity", solid_damping="solidDamping",
                  floor_friction_solid="floorFrictionSolid", floor_friction_fluid="floorFrictionFluid",
                  cull_radius="cullRadius", max_speed="maxSpeed", grab_stiffness="grabStiffness",
                  cell_size="cellSize")


def _scene_stats(sc, name, frames):
    """A GPU-scale scene as the reference runs it: the scene's constants set
    on the solver's public fields (cellSize after addSolidCube, which
    allocates), its emits, its wedges within `frames`; statistics only."""
    p = sc.params_with()
    assert p["fluid_damping"] == 0.6
    fields = {js: float(p[k]) for k, js in _JS_FIELDS.items()}
    return dict(
        name=name, maxFluid=sc.max_fluid, gravity=float(p["gravity"]), cube=[float(x) for x in sc.cube],
        fields=fields,
        emit=[[*map(float, q), 0.0, 0.0, 0.0, sc.fluid_mass] for q in sc.fluid_pos],
        frames=frames, substeps=sc.substeps,
        cuts=[dict(frame=w.frame, o=list(w.o), d1=list(w.d1), d2=list(w.d2)) for w in sc.wedges if w.frame < frames],
        dumpFrames=[], dumpInit=False, statsOnly=True,
    )


def single_gpu_stats():
    """The 20^3 + 65,536 scene (its wedge at frame 150) for 300 frames."""
    return _scene_stats(scenes.single_gpu_scene(), "single_gpu_stats", 300)


def main():
    out = os.path.join(HERE, "inputs")
    os.makedirs(out, exist_ok=True)
    for d in (plain_cpu(), grab_tear(), ring_cap(), single_gpu_stats()):
        with open(os.path.join(out, d["name"] + ".json"), "w") as f:
            json.dump(d, f)
        print("wrote", d["name"], len(d["emit"]), "emits")


if __name__ == "__main__":
    main()
