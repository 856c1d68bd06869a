"""Edge-shaped scenes against the FP64 oracle: the smallest cube (one solid,
no constraints) beside fluid, a cube with no fluid at all (maxFluid 0), and
fluid emitted outside the cull radius (src/solver.js:423-425 deactivates it
in the first substep), plus an empty emit.

Bar as in test_gpu_parity: after a few frames, one substep within 1e-4 * h
of the FP64 oracle (Jacobi form) from the same state, active flags equal;
the fluid count equals the oracle's.
"""
import numpy as np
import pytest

from tests._gpu import log
from tests.test_gpu_rows import _one_substep_vs_oracle

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


@pytest.fixture(scope="module", autouse=True)
def _lib():
    from fluidphysics_amd import _abi
    _abi.load_library()


def _block(count, x0, culled=0):
    """count fluid particles on an 8 x 8 lattice at the slab spacing from x0;
    the first `culled` of them moved 8 m out (beyond the 7 m cull radius)."""
    k = np.arange(count)
    pos = np.stack([x0 + (k % 8) * 0.041, 0.05 + (k // 64) * 0.041, ((k // 8) % 8) * 0.041 - 0.14], 1)
    pos = pos.astype(np.float64)
    pos[:culled, 0] += 8.0
    return pos


def _edge_scene(name):
    from fluidphysics_amd.scenes import Scene, gpu_scale_params
    r = 0.03
    if name == "single_solid":
        n, fluid = 1, _block(512, 0.3)
    elif name == "cube_no_fluid":
        n, fluid = 3, np.zeros((0, 3))
    else:  # culled_fluid
        n, fluid = 2, _block(128, 0.3, culled=64)
    return Scene(name=name, params=gpu_scale_params(r), max_fluid=len(fluid), cube=(0.0, 0.5, 0.0, n, 1.91 * r, 100 * r ** 3),
                 fluid_pos=fluid, fluid_vel=np.zeros_like(fluid), fluid_mass=12 * r ** 3, frames=8)


@pytest.mark.parametrize("name", ["single_solid", "cube_no_fluid", "culled_fluid"])
def test_edge_scene_one_substep_vs_fp64(name):
    from fluidphysics_amd import from_scene
    from oracle.oracle import OracleSolver
    sc = _edge_scene(name)
    log("case edge scene %s: cube %d^3, %d fluid" % (name, sc.cube[3], sc.max_fluid))
    g = from_scene(sc)
    g.emit_batch(np.zeros((0, 3)), np.zeros((0, 3)), 0.01)  # an empty emit is a no-op
    o = OracleSolver(sc.params_with(), sc.max_fluid, "f64", "jacobi")
    o.add_solid_cube(*sc.cube)
    if sc.max_fluid:
        o.emit(sc.fluid_pos, sc.fluid_vel, sc.fluid_mass)
    for _ in range(6):
        g.step(DT, sc.substeps)
        o.step(DT, sc.substeps)
    assert g.fluidCount() == int(o.read_state()["active"][g.solidCount:].sum())
    if name == "culled_fluid":
        assert g.fluidCount() == 64
    assert not g.nonfinite()
    _one_substep_vs_oracle(g, sc, name)
    g.close()
