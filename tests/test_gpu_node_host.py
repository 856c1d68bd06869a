"""The Node.js host (fluidphysics_amd/js/solver.mjs over the N-API addon)
drives the same HIP kernels: identical inputs give bit-identical results to
the ctypes host, the plain scene's first substep is within 1e-4*h of the
FP64 replay of the reference, and the headless version of the reference's
browser drive test (test/drive.js timeline) runs clean."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests import _golden as G
from tests._gpu import log, make_gpu

pytestmark = pytest.mark.gpu

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
NODE = shutil.which("node")
ADDON = os.path.join(ROOT, "fluidphysics_amd", "addon", "fp_napi.node")


def _need_node():
    if NODE is None:
        pytest.skip("node is not installed")
    assert os.path.exists(ADDON), "fp_napi.node is not built (run __graft_entry__.build())"


def _run_node(d, tmp_path, frames, sub, dt):
    inp = tmp_path / "in.json"
    out = tmp_path / "out.bin"
    inp.write_text(json.dumps(d))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "run_scene.mjs"), str(inp), str(out),
                        str(frames), str(sub), repr(dt)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    n = info["capacity"]
    raw = out.read_bytes()
    pos = np.frombuffer(raw, np.float32, 3 * n, 0)
    vel = np.frombuffer(raw, np.float32, 3 * n, 12 * n)
    act = np.frombuffer(raw, np.uint8, n, 24 * n)
    return pos, vel, act


def test_node_host_bit_identical_to_ctypes_host(tmp_path):
    _need_node()
    log("case node host vs ctypes host plain_cpu 20 frames")
    d = G.scenario("plain_cpu")
    d["params"] = {"constraintIters": 3, "jacobiOmega": 8.0}
    pos, vel, act = _run_node(d, tmp_path, 20, 5, 1.0 / 60.0)
    g = make_gpu(d, iters=3, omega=8.0)
    drv = G.Driver(d)
    for f in range(20):
        for c in d["cuts"]:
            if c["frame"] == f:
                g.cut(*c["o"], *c["d1"], *c["d2"])
        g.step(1.0 / 60.0, 5)
    st = g.read_state()
    assert np.array_equal(st["pos"], pos)
    assert np.array_equal(st["vel"], vel)
    assert np.array_equal(st["active"], act)


def test_node_host_solid_modes_bit_identical_to_ctypes_host(tmp_path):
    """The optional solid modes set through the Node host's tuning names
    (tetVolume, solidContactCache): 110 frames with the slice at frame 100
    equal the ctypes host's bit for bit, tets included."""
    _need_node()
    log("case node host vs ctypes host, solid modes, plain_cpu 110 frames")
    d = G.scenario("plain_cpu")
    d["params"] = {"tetVolume": 1, "solidContactCache": 1}
    d["dumpTets"] = True
    pos, vel, act = _run_node(d, tmp_path, 110, 5, 1.0 / 60.0)
    raw = (tmp_path / "out.bin.tets").read_bytes()
    g = make_gpu(d, iters=None, omega=None, tet_volume=1, solid_contact_cache=1)
    for f in range(110):
        for c in d["cuts"]:
            if c["frame"] == f:
                g.cut(*c["o"], *c["d1"], *c["d2"])
        g.step(1.0 / 60.0, 5)
    st = g.read_state()
    v, a = g.read_tets()
    nt = len(a)
    assert nt == 5 * 6 ** 3 and len(raw) == 17 * nt
    assert np.array_equal(np.frombuffer(raw, np.int32, 4 * nt, 0).reshape(-1, 4), v)
    assert np.array_equal(np.frombuffer(raw, np.uint8, nt, 16 * nt), a)
    assert a.sum() < nt  # the slice killed tets
    assert np.array_equal(st["pos"], pos)
    assert np.array_equal(st["vel"], vel)
    assert np.array_equal(st["active"], act)


def test_node_host_first_substep_vs_fp64_reference(tmp_path):
    _need_node()
    log("case node host one substep plain_cpu")
    d = G.scenario("plain_cpu")
    d["params"] = {"constraintIters": 3, "jacobiOmega": 8.0}
    sub = 1.0 / 60.0 / 5
    pos, vel, act = _run_node(d, tmp_path, 1, 1, sub)
    o = G.make_oracle(d, storage="f64", mode="gs", iters=1)
    o.step(sub, 1)
    ref = o.read_state64()
    m = np.repeat(act.astype(bool), 3)
    assert np.max(np.abs(pos[m] - ref["pos"][m])) <= 1e-4 * 0.19
    assert np.max(np.abs(vel[m] - ref["vel"][m])) * sub <= 1e-4 * 0.19


def test_headless_drive_timeline():
    _need_node()
    log("case node drive timeline")
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "drive.mjs")], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr
    phases = [json.loads(l) for l in r.stdout.strip().splitlines()]
    log("  " + " | ".join("%s f=%d fluid=%d alive=%d ke=%.2f" % (p["phase"], p["frames"], p["fluid"], p["alive"], p["ke"])
                          for p in phases))
    by = {p["phase"]: p for p in phases}
    assert [p["phase"] for p in phases] == ["settle", "slice", "grab", "hose"]
    assert all(p["nan"] == 0 and p["nonFinite"] is False for p in phases)
    assert by["slice"]["alive"] < by["settle"]["alive"]      # the strokes cut constraints
    assert by["hose"]["fluid"] > 0                            # the hose emitted water
    assert by["settle"]["minY"] >= 0.09 - 1e-3                # the cube rests on the floor
    # the device components (components()) partition the solids as `group` does
    assert all(p["same"] for p in phases), [(p["phase"], p["comps"]) for p in phases]
    assert by["settle"]["comps"] == 1


def to_js_params(p):
    """fp_params field names (snake_case) -> the addon's camelCase names."""
    return {k.split("_")[0] + "".join(w.title() for w in k.split("_")[1:]): v for k, v in p.items()}


def test_node_batch_bit_identical_to_ctypes_batch(tmp_path):
    """The Node host's Batch (solver.mjs over the addon) and the ctypes
    host's Batch step the same scenes to bit-identical snapshots, and the
    addon's sceneBytes equals what each solver holds."""
    _need_node()
    from fluidphysics_amd import Batch, from_scene, scenes
    frames, sub, dt = 6, 5, 1.0 / 60.0
    scs = [scenes.batched_scene(k, frames=frames) for k in (50, 51, 52)]
    for sc in scs:
        for j, w in enumerate(sc.wedges):
            w.frame = 2 + j
    log("case node batch vs ctypes batch: %s" % [sc.name for sc in scs])
    doc = []
    for sc in scs:
        doc.append(dict(params=to_js_params(sc.params_with()), maxFluid=sc.max_fluid, cube=list(sc.cube),
                        pos=sc.fluid_pos.ravel().tolist(), vel=sc.fluid_vel.ravel().tolist(), mass=sc.fluid_mass,
                        cuts=[dict(frame=w.frame, o=list(w.o), d1=list(w.d1), d2=list(w.d2)) for w in sc.wedges]))
    inp = tmp_path / "scenes.json"
    out = tmp_path / "snap.bin"
    inp.write_text(json.dumps(doc))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "run_batch.mjs"), str(inp), str(out),
                        str(frames), str(sub), repr(dt)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    for est, held in res["bytes"]:
        assert est == held
    total = res["offsets"][-1]
    raw = np.frombuffer(out.read_bytes(), np.float32).reshape(frames, total)
    b = Batch()
    ss = [from_scene(sc) for sc in scs]
    for s in ss:
        b.add(s)
    for f in range(frames):
        for s, sc in zip(ss, scs):
            for w in sc.wedges:
                if w.frame == f:
                    s.cut(*w.o, *w.d1, *w.d2)
        b.step(dt, sub)
        flat, off = b.snapshot()
        assert list(off) == res["offsets"]
        assert np.array_equal(flat, raw[f]), "frame %d" % f
    assert res["info"]["graphLaunches"] == frames - 4  # after the contact-lane tuning frames
    b.close()
