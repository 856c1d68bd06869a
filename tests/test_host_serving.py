"""Host side of the batched and serving workloads (no device): the request
stream and scene generators are deterministic and follow their stated
distributions; the admission estimate needs no device."""
import math

import numpy as np
import pytest

from fluidphysics_amd import scenes


def test_batched_scene_distribution_and_determinism():
    a = scenes.batched_scenes(64, seed=4, frames=600)
    b = scenes.batched_scenes(64, seed=4, frames=600)
    for x, y in zip(a, b):
        assert x.name == y.name and np.array_equal(x.fluid_pos, y.fluid_pos)
        assert x.params == y.params
    dims = np.array([s.cube[3] for s in a])
    nf = np.array([s.max_fluid for s in a])
    assert dims.min() >= 6 and dims.max() <= 20
    assert nf.min() >= 1000 and nf.max() <= 65536
    for s in a:
        g = s.params
        tilt = math.degrees(math.atan2(math.hypot(g["gravity_x"], g["gravity_z"]), -g["gravity"]))
        assert 0.0 <= tilt <= 15.0 + 1e-9
        assert abs(math.sqrt(g["gravity"] ** 2 + g["gravity_x"] ** 2 + g["gravity_z"] ** 2) - 9.8) < 1e-9
        assert len(s.fluid_pos) == s.max_fluid
        assert 1 <= len(s.wedges) <= 2
    # a scene does not depend on how many others are generated
    assert scenes.batched_scene(5, 4).name == a[5].name


def test_serving_requests_poisson_stream():
    r = scenes.serving_requests(200.0, 20.0, seed=5)
    assert r == scenes.serving_requests(200.0, 20.0, seed=5)
    t = np.array([x.arrival for x in r])
    assert np.all(np.diff(t) > 0) and t[-1] < 20.0
    # rate within 4 sigma of 200/s; horizons log-uniform over 1..600
    assert abs(len(r) - 4000) < 4 * math.sqrt(4000)
    h = np.array([x.horizon for x in r])
    assert h.min() >= 1 and h.max() <= 600
    assert abs(np.median(np.log(h)) - 0.5 * math.log(600)) < 0.15


def test_scene_bytes_without_device():
    from fluidphysics_amd import scene_bytes
    p = scenes.gpu_scale_params()
    a = scene_bytes(p, 4000, 10)
    assert a > 0
    # sizes are rounded up to the device pool's size classes: a little more
    # never takes less, a lot more takes more
    assert scene_bytes(p, 4000, 12) >= a and scene_bytes(p, 8000, 10) > a and scene_bytes(p, 4000, 20) > a
    assert scene_bytes(p, 0, 6) > 0
    with pytest.raises(Exception):
        scene_bytes(p, 100, 0)
