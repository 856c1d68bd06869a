"""CPU anchor of the measurement (TEST INFRASTRUCTURE; development container
only): the reference CPU solver src/solver.js under node, and this repo's C
restatement of it (the oracle, Gauss-Seidel mode, float32 stores -- the
bench's cpu_baseline "port"), timed single-threaded on the same frames of the
three single-scene workloads, from their initial states.

    python3 oracle/time_reference.py [--repeats 3]

Prints one JSON line per workload: median ms per frame of each, particle
substeps per second, and the port : solver.js speed ratio.  node runs
oracle/time_reference.mjs, which imports /root/reference/src/solver.js.
"""
import argparse
import importlib.util
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

from fluidphysics_amd import scenes  # noqa: E402
from oracle.oracle import OracleSolver  # noqa: E402


def _gen():
    spec = importlib.util.spec_from_file_location("gen_inputs", os.path.join(ROOT, "tests", "golden", "gen_inputs.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def workloads():
    g = _gen()
    plain = g.plain_cpu()
    plain["fields"] = {}
    out = [("plain_cpu", scenes.plain_cpu_scene(), plain, 60),
           ("single_gpu", scenes.single_gpu_scene(), g._scene_stats(scenes.single_gpu_scene(), "single_gpu", 10), 10)]
    large = scenes.large_scene()
    d = g._scene_stats(large, "large", 3)
    sp = 1.91 * large.params_with()["fluid_radius"]
    d["hose"] = []
    for f in range(3):
        p, v = scenes.hose_burst(large.hose, f, sp)
        d["hose"].append([[*map(float, a), *map(float, b), large.fluid_mass] for a, b in zip(p, v)])
    out.append(("large", large, d, 3))
    return out


def time_port(sc, d, frames):
    o = OracleSolver(sc.params_with(constraint_iters=1), sc.max_fluid, "f32", "gs")
    o.add_solid_cube(*sc.cube)
    e = np.asarray(d["emit"], np.float64)
    o.emit(e[:, 0:3], e[:, 3:6], e[:, 6])
    t0 = time.perf_counter()
    for f in range(frames):
        h = np.asarray(d.get("hose", [[]] * frames)[f], np.float64).reshape(-1, 7)
        if len(h):
            o.emit(h[:, 0:3], h[:, 3:6], h[:, 6])
        o.step(1.0 / 60.0, d["substeps"])
    ms = (time.perf_counter() - t0) * 1e3 / frames
    return ms, sc.solid_count + o.fluid_count()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--repeats", type=int, default=3)
    args = ap.parse_args()
    cores = os.cpu_count()
    for name, sc, d, frames in workloads():
        with tempfile.NamedTemporaryFile("w", suffix=".json", delete=False) as f:
            json.dump(d, f)
            path = f.name
        try:
            r = subprocess.run(["node", os.path.join(HERE, "time_reference.mjs"), path, str(frames),
                                str(args.repeats)], capture_output=True, text=True, check=True)
        finally:
            os.unlink(path)
        js = json.loads(r.stdout.strip().splitlines()[-1])
        port = [time_port(sc, d, frames) for _ in range(args.repeats)]
        js_ms = float(np.median(js["ms_per_frame"]))
        port_ms = float(np.median([p[0] for p in port]))
        act = js["active"]
        assert act == port[0][1], (act, port[0][1])
        print(json.dumps({
            "workload": name, "frames": frames, "repeats": args.repeats, "active_particles": act,
            "solver_js_ms_per_frame": round(js_ms, 2), "solver_js_ms_all": [round(x, 2) for x in js["ms_per_frame"]],
            "port_ms_per_frame": round(port_ms, 2),
            "solver_js_particle_substeps_per_s": round(act * d["substeps"] / js_ms * 1e3, 1),
            "port_particle_substeps_per_s": round(act * d["substeps"] / port_ms * 1e3, 1),
            "port_over_solver_js_speed": round(js_ms / port_ms, 3),
            "threads": 1, "host_cores": cores, "node": subprocess.run(["node", "--version"], capture_output=True,
                                                                       text=True).stdout.strip(),
        }), flush=True)


if __name__ == "__main__":
    main()
