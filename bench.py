"""Benchmark: particle-substeps/s of the unified-particle solver on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload large]

A step is one frame of the workload: the hose burst, any scheduled slice /
grab, then `substeps` solver substeps (src/solver.js:187 step(dt, 5)).
Inputs are resident in HBM when the timed region starts; the region is
bracketed by a device sync on both sides.  value = (active particles x
substeps x frames) / wall time.

Workloads (fluidphysics_amd/scenes.py):
  large       24^3 = 13,824 cube + 262,144 fluid slab + 120/frame hose (default;
              the configuration the throughput target is quoted on)
  single_gpu  20^3 = 8,000 cube + 65,536 fluid column
  plain_cpu   7^3 = 343 cube + 2,600 fluid (the reference's CPU scale)

A second leg steps the batched workload (256 independent scenes of uneven
sizes as one batch, fluidphysics_amd.scenes.batched_scene) and reports its
aggregate throughput and frame p50/p99 beside the same scenes stepped one at
a time (the `batched` object; --no-batched skips it).  A third serves an
open-loop Poisson stream of scene requests with continuous batching under an
HBM budget (fluidphysics_amd.serving) and reports frame p50/p99 and
sustained throughput beside the per-scene path on the same stream (the
`serving` object; --no-serving skips it).  At the default 120 requests/s the
batched server keeps up with the stream, so its throughput is the stream's
load; `serving.capacity` repeats both paths on a heavier stream (240/s by
default) to show each one's throughput where it, not the stream, is the
limit, with the batched path's frame and queueing latency beside it.

Printed: one JSON line with the metric, a `roofline` object for the dominant
kernel and a `cpu_baseline` object.  The roofline's launch time is measured
over the timed frames themselves: the state at the start of the timed region
is kept in a device checkpoint, and after the timed run the same frames are
replayed from it with HIP events around every pass on the solver's stream
(the replay is bit-identical to the timed run, which the line reports).
cpu_baseline is the oracle -- the plain-C restatement of src/solver.js,
Gauss-Seidel mode, float32 stores -- on a bounded sample of the same
workload: one host core, and one independent copy per core of the box's CPU
share (aggregate), both run before the GPU is touched.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (scaled by a plain copy kernel of known size); see the file for commands.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "large_hbm_traffic.json")
# Speed of the oracle port over the reference src/solver.js under node, same
# frames, one thread each (oracle/time_reference.py in the development
# container; profiles/cpu_reference_timing.md): a port rate divided by this
# is the reference's own rate on the same host.
PORT_OVER_SOLVER_JS = {"plain_cpu": 1.69, "single_gpu": 1.86, "large": 1.85}


def model_bytes_per_substep(n, nf, ns, c, t, iters):
    """Algorithmic bytes per substep: each array a pass reads or writes counted
    once at the float4 SoA layout (integrate 84, grid count 28, scatter 32,
    contacts 44, bounds 56, velocity 52 per particle; DDR density 36,
    displacement 52, XSPH 60 per fluid particle; 12 per table entry; per
    constraint iteration 40 per constraint + 36 per solid)."""
    return 296 * n + 148 * nf + 12 * t + iters * (40 * c + 36 * ns)


# Per-launch algorithmic bytes of each pass under the same model (per-unit
# figure x the units one launch processes).
def pass_bytes(name, n, nf, ns, c, t):
    return {
        "integrate": 84 * n,
        "hash_count": 28 * n,
        "scatter": 32 * n,
        "contacts": 44 * n,
        "density": 36 * nf,
        "displace": 52 * nf,
        "constraints": 40 * c + 36 * ns,
        "bounds_velocity": (56 + 52) * n,
        "xsph": 60 * nf,
        "scan_reduce": 4 * t,
        "scan_apply": 8 * t,
    }.get(name)


def measured_traffic(workload, pass_name):
    """HBM bytes per launch of the pass's kernel on this workload, from the
    committed rocprofv3 counter summary (None when absent)."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload:
        return None
    ks = t.get("kernels", {})
    # the pass's kernel, or its one template instance in the run (kg_density<false>)
    k = ks.get("kg_" + pass_name)
    if k is None:
        inst = [v for n, v in ks.items() if n.startswith("kg_" + pass_name + "<")]
        k = inst[0] if len(inst) == 1 else None
    return int(k["hbm_bytes"]) if k else None


def run_frames(solver, scene, f0, frames, rest_pos, grab_ids, sync_each=False):
    from fluidphysics_amd import scenes
    sp = 1.91 * scene.params_with()["fluid_radius"]
    for f in range(f0, f0 + frames):
        if scene.hose:
            p, v = scenes.hose_burst(scene.hose, f, sp)
            solver.emit_batch(p, v, scene.fluid_mass)
        for w in scene.wedges:
            if w.frame == f:
                solver.cut(*w.o, *w.d1, *w.d2)
        if scene.grab and grab_ids is not None:
            g0, g1 = scene.grab["frames"]
            if g0 <= f < g1:
                tg = scenes.grab_targets(scene, f, grab_ids, rest_pos)
                solver.grabs = [{"i": int(i), "x": float(x[0]), "y": float(x[1]), "z": float(x[2])}
                                for i, x in zip(grab_ids, tg)]
            elif f == g1:
                solver.grabs = []
        solver.step(scene.dt, scene.substeps, wait=sync_each)


def active_count(solver):
    return solver.solidCount + solver.fluidCount()


def cpu_baseline(scene, budget_s=12.0, max_frames=50):
    """The oracle (plain-C restatement of src/solver.js, Gauss-Seidel mode,
    float32 stores) on one host core: frames of the same workload from its
    initial state until ~budget_s of CPU work."""
    from oracle.oracle import OracleSolver
    from fluidphysics_amd import scenes
    p = scene.params_with(constraint_iters=1)
    o = OracleSolver(p, scene.max_fluid, storage="f32", mode="gs")
    o.add_solid_cube(*scene.cube)
    o.emit(scene.fluid_pos, scene.fluid_vel, scene.fluid_mass)
    sp = 1.91 * p["fluid_radius"]
    psub = 0
    frames = 0
    t0 = time.perf_counter()
    while frames < max_frames and time.perf_counter() - t0 < budget_s:
        if scene.hose:
            hp, hv = scenes.hose_burst(scene.hose, frames, sp)
            o.emit(hp, hv, scene.fluid_mass)
        n_act = scene.solid_count + o.fluid_count()
        o.step(scene.dt, scene.substeps)
        psub += n_act * scene.substeps
        frames += 1
    el = time.perf_counter() - t0
    return {"value": psub / el, "unit": "particle-substeps/s", "cores": 1, "kind": "port",
            "sample": "%d frame(s) x %d substeps of the %s workload from its initial state (%.1f s), "
                      "oracle Gauss-Seidel mode (src/solver.js semantics), float32 stores, 1 thread of %d"
                      % (frames, scene.substeps, scene.name, el, os.cpu_count() or 0)}


def _cpu_worker(args):
    name, budget_s, max_frames = args
    from fluidphysics_amd import scenes
    return cpu_baseline(scenes.SCENES[name](), budget_s, max_frames)


def host_cores():
    """The CPU share this process may use (affinity), at most 16 (the share
    of one GPU on the box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline_all_cores(workload, budget_s=10.0, max_frames=50):
    """One independent copy of the same sample per host core, run together
    (the reference solver is single-threaded; this is the host's aggregate
    rate on independent scenes).  Forked before the GPU is initialised."""
    import multiprocessing as mp
    n = host_cores()
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(n) as pool:
        outs = pool.map(_cpu_worker, [(workload, budget_s, max_frames)] * n)
    el = time.perf_counter() - t0
    return {"value": round(sum(o["value"] for o in outs), 1), "unit": "particle-substeps/s", "cores": n,
            "kind": "port", "processes": n, "wall_s": round(el, 2),
            "sample": "%d independent copies of the 1-core sample, one per core, run at once (aggregate)" % n}


def _pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(math.ceil(q * len(xs))) - 1)]


def batched_leg(count, frames, warmup, solo_frames, seed=4):
    """The batched workload: `count` independent scenes (scenes.batched_scene:
    cube 6^3..20^3, 1k..64k fluid, tilted gravity, slice schedules) stepped
    as one Batch -- one launch per pass over every scene, the frame replayed
    from a HIP graph -- and then the same scenes stepped one at a time
    through their own solvers.  Frame latency = one frame of every scene,
    slices included, measured host-side around a completed frame."""
    from fluidphysics_amd import Batch, from_scene, scenes
    horizon = warmup + frames + solo_frames + 1
    t0 = time.perf_counter()
    scs = scenes.batched_scenes(count, seed, frames=horizon)
    ss = [from_scene(sc) for sc in scs]
    setup_s = time.perf_counter() - t0
    dt, sub = scs[0].dt, scs[0].substeps

    def cuts(f):
        for s, sc in zip(ss, scs):
            for w in sc.wedges:
                if w.frame == f:
                    s.cut(*w.o, *w.d1, *w.d2)

    def active():
        return sum(active_count(s) for s in ss)

    b = Batch()
    for s in ss:
        b.add(s)
    for f in range(warmup):
        cuts(f)
        b.step(dt, sub)
    n0 = active()
    for s in ss:
        s.checkpoint(0)
    lat = []
    for f in range(warmup, warmup + frames):
        t = time.perf_counter()
        cuts(f)
        b.step(dt, sub)
        lat.append(time.perf_counter() - t)
    n1 = active()
    slow = sorted(range(len(lat)), key=lambda k: -lat[k])[:3]
    print("bench: batched slowest frames %s" % ", ".join("%d: %.2f ms" % (warmup + k, lat[k] * 1e3) for k in slow),
          file=sys.stderr, flush=True)
    inf = b.info()
    n_mean = 0.5 * (n0 + n1)
    value = n_mean * sub * frames / sum(lat)
    # the timed frames again from the checkpoints, with events around every pass
    for s in ss:
        s.restore(0)
    b.profile(True)
    for f in range(warmup, warmup + frames):
        cuts(f)
        b.step(dt, sub)
    b.sync()
    stats = b.profile_read()
    b.profile(False)
    ns_t = sum(s.solidCount for s in ss)
    infos = [s.info() for s in ss]
    roof, _ = roofline_of(stats, frames, n_mean, n_mean - ns_t, ns_t, sum(i.num_constraints for i in infos),
                          sum(i.table_size for i in infos), first_frame=warmup)
    for s in ss:
        b.remove(s)
    b.close()

    solo = None
    if solo_frames > 0:
        f0 = warmup + frames
        for s in ss:  # each solver captures its own frame graph before timing
            s.step(dt, sub, wait=False)
        for s in ss:
            s.sync()
        lat1 = []
        for f in range(f0 + 1, f0 + 1 + solo_frames):
            t = time.perf_counter()
            cuts(f)
            for s in ss:
                s.step(dt, sub, wait=False)
            for s in ss:
                s.sync()
            lat1.append(time.perf_counter() - t)
        n2 = active()
        value1 = 0.5 * (n1 + n2) * sub * solo_frames / sum(lat1)
        solo = {"value": round(value1, 1), "frames": solo_frames,
                "frame_ms_p50": round(_pct(lat1, 0.50) * 1e3, 3),
                "frame_ms_p99": round(_pct(lat1, 0.99) * 1e3, 3)}
    for s in ss:
        s.close()
    return {
        "workload": "batched_%d_seed%d" % (count, seed),
        "scenes": count,
        "solids": int(sum(sc.solid_count for sc in scs)),
        "fluid": int(sum(sc.max_fluid for sc in scs)),
        "active_mean": int(n_mean),
        "frames": frames,
        "value": round(value, 1),
        "unit": "particle-substeps/s",
        "frame_ms_p50": round(_pct(lat, 0.50) * 1e3, 3),
        "frame_ms_p99": round(_pct(lat, 0.99) * 1e3, 3),
        "graph_launches": int(inf.graph_launches),
        "roofline": roof,
        "device_bytes": int(inf.device_bytes),
        "setup_s": round(setup_s, 2),
        "one_at_a_time": solo,
        "speedup_vs_one_at_a_time": round(value / solo["value"], 3) if solo else None,
    }


def serving_leg(rate_hz, duration_s, budget_bytes, seed=5):
    """The serving workload: an open-loop Poisson stream of scene requests
    (rate_hz, seeded; sizes as in the batched workload, horizons
    log-uniform over 1..600 frames) served for duration_s seconds by the
    continuous-batching loop (fluidphysics_amd.serving), then the same
    stream through the plain per-scene path.  Frame latency = dispatch of a
    frame -> its position snapshot on the host."""
    from fluidphysics_amd import scenes
    from fluidphysics_amd.serving import Server
    reqs = scenes.serving_requests(rate_hz, duration_s, seed=seed)
    out = {"workload": "serving_poisson_%gHz_seed%d" % (rate_hz, seed), "rate_hz": rate_hz,
           "duration_s": duration_s, "requests": len(reqs)}
    for mode in ("batch", "per_scene"):
        rep = Server(budget_bytes=budget_bytes, mode=mode).run(reqs, duration=duration_s)
        out[mode] = rep.summary()
    out["speedup_vs_per_scene"] = round(out["batch"]["value"] / max(out["per_scene"]["value"], 1.0), 3)
    return out


def roofline_of(stats, frames, n_mean, nf, ns, c, t, traffic=None, first_frame=None):
    """The dominant pass of per-pass event statistics: algorithmic bytes per
    launch / average launch time, against HBM's 8 TB/s, plus every pass's
    figures."""
    passes = {}
    dom, dom_ms = None, -1.0
    for name, (launches, total_ms) in stats.items():
        avg = total_ms / launches
        passes[name] = {"launches": launches, "avg_us": round(avg * 1e3, 3),
                        "ms_per_frame": round(total_ms / frames, 4)}
        b = pass_bytes(name, n_mean, nf, ns, c, t)
        if b is not None:
            passes[name]["gbs"] = round(b / (avg * 1e-3) / 1e9, 1)
            if total_ms > dom_ms:
                dom, dom_ms = name, total_ms
    launches, total_ms = stats[dom]
    avg_s = total_ms / launches * 1e-3
    bytes_dom = pass_bytes(dom, n_mean, nf, ns, c, t)
    achieved = bytes_dom / avg_s / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": "kg_" + dom,
            "bytes_per_launch": int(bytes_dom), "avg_launch_us": round(avg_s * 1e6, 3), "launches": launches,
            "timed_frames": [first_frame, first_frame + frames] if first_frame is not None else None,
            "timing": "HIP events around the pass on its stream, over a replay of the timed frames from a "
                      "device checkpoint taken at their start (eager launches)"}
    if traffic:
        roof["traffic_gbs"] = round(traffic / avg_s / 1e9, 1)
    return roof, passes


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="large", choices=["large", "single_gpu", "plain_cpu"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched-scenes", type=int, default=256)
    ap.add_argument("--batched-steps", type=int, default=60)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-solo-frames", type=int, default=-1,
                    help="frames of the one-at-a-time comparison (-1: a third of --batched-steps, 0: skip)")
    ap.add_argument("--serving-rate", type=float, default=120.0)
    ap.add_argument("--serving-capacity-rate", type=float, default=240.0,
                    help="a second, heavier stream (0: skip): each path's throughput where it is the bottleneck")
    ap.add_argument("--serving-seconds", type=float, default=10.0)
    ap.add_argument("--serving-budget-gb", type=float, default=200.0)
    ap.add_argument("--no-serving", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the solver's path does not span GPUs "
              "(multi-GPU sharding is a design note only). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from fluidphysics_amd import from_scene, scenes
    scene = scenes.SCENES[args.workload]()

    # the CPU baselines first: the all-cores leg forks workers, which must
    # happen before this process initialises the GPU
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scene)
        cpu["all_cores"] = cpu_baseline_all_cores(args.workload)
        cpu["port_over_solver_js"] = PORT_OVER_SOLVER_JS.get(args.workload)
        print("bench: cpu baseline %.3e p-substeps/s on 1 core, %.3e on %d" %
              (cpu["value"], cpu["all_cores"]["value"], cpu["all_cores"]["cores"]), file=sys.stderr, flush=True)

    solver = from_scene(scene)
    inf = solver.info()
    p = scene.params_with()
    grab_ids = rest_pos = None
    if scene.grab:
        st = solver.read_state()
        pos = st["pos"].reshape(-1, 3)[: solver.solidCount]
        d = np.linalg.norm(pos - np.asarray(scene.grab["point"]), axis=1)
        grab_ids = np.argsort(d, kind="stable")[: scene.grab["count"]]
        rest_pos = pos[grab_ids].astype(np.float64)

    print("bench: workload %s, capacity %d, table %d, constraints %d" %
          (scene.name, inf.capacity, inf.table_size, inf.num_constraints), file=sys.stderr, flush=True)
    run_frames(solver, scene, 0, args.warmup, rest_pos, grab_ids)
    solver.sync()
    n0 = active_count(solver)
    f0 = args.warmup
    solver.checkpoint(0)   # the state the timed frames start from (outside the timed region)
    solver.sync()
    t0 = time.perf_counter()
    run_frames(solver, scene, f0, args.steps, rest_pos, grab_ids)
    solver.sync()
    el = time.perf_counter() - t0
    n1 = active_count(solver)
    n_mean = 0.5 * (n0 + n1)
    value = n_mean * scene.substeps * args.steps / el
    end_pos = solver.read_positions().copy()
    print("bench: %d frames in %.3f s, %.0f active particles" % (args.steps, el, n_mean), file=sys.stderr, flush=True)

    # the timed frames again from the checkpoint, with events around every pass
    solver.restore(0)
    solver.profile(True)
    run_frames(solver, scene, f0, args.steps, rest_pos, grab_ids)
    solver.sync()
    stats = solver.profile_read()
    solver.profile(False)
    replay_identical = bool(np.array_equal(end_pos, solver.read_positions()))

    nf = n_mean - solver.solidCount
    ns, c, t = solver.solidCount, inf.num_constraints, inf.table_size
    iters = p["constraint_iters"]
    roofline, passes = roofline_of(stats, args.steps, n_mean, nf, ns, c, t, first_frame=f0)
    roofline["traffic"] = measured_traffic(scene.name, roofline["kernel"][3:])
    if roofline["traffic"]:
        roofline["traffic_gbs"] = round(roofline["traffic"] / (roofline["avg_launch_us"] * 1e-6) / 1e9, 1)
    roofline["replay_identical"] = replay_identical
    inf_end = solver.info()
    b_sub = model_bytes_per_substep(n_mean, nf, ns, c, t, iters)
    model_frac = b_sub * (value / n_mean) / (HBM_PEAK_GBS * 1e9)
    solver.close()

    batched = None
    if not args.no_batched:
        solo = args.batched_solo_frames if args.batched_solo_frames >= 0 else max(10, args.batched_steps // 3)
        # warm-up covers the one-time frames: the contact-lane tuning frames,
        # the launch-shape tuning frames and the frames that capture graphs
        batched = batched_leg(args.batched_scenes, args.batched_steps, max(20, args.warmup // 2), solo)
        print("bench: batched %.3e p-substeps/s, frame p50 %.2f ms p99 %.2f ms; one at a time %s" %
              (batched["value"], batched["frame_ms_p50"], batched["frame_ms_p99"],
               batched["one_at_a_time"] and "%.3e" % batched["one_at_a_time"]["value"]), file=sys.stderr, flush=True)

    serving = None
    if not args.no_serving:
        serving = serving_leg(args.serving_rate, args.serving_seconds, args.serving_budget_gb * 1e9)
        if args.serving_capacity_rate > 0:
            serving["capacity"] = serving_leg(args.serving_capacity_rate, args.serving_seconds,
                                              args.serving_budget_gb * 1e9)
            cap = serving["capacity"]
            print("bench: serving at %g Hz: %.3e p-substeps/s, frame p99 %.2f ms, queue p99 %.2f ms; per-scene %.3e" %
                  (args.serving_capacity_rate, cap["batch"]["value"], cap["batch"]["frame_ms_p99"],
                   cap["batch"]["queue_ms_p99"], cap["per_scene"]["value"]), file=sys.stderr, flush=True)
        print("bench: serving %.3e p-substeps/s, frame p50 %.2f ms p99 %.2f ms; per-scene %.3e p99 %.2f ms" %
              (serving["batch"]["value"], serving["batch"]["frame_ms_p50"], serving["batch"]["frame_ms_p99"],
               serving["per_scene"]["value"], serving["per_scene"]["frame_ms_p99"]), file=sys.stderr, flush=True)

    out = {
        "metric": "particle-substeps/s",
        "value": round(value, 1),
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": scene.name, "solids": ns, "fluid_active_mean": int(nf),
                   "capacity": inf.capacity, "constraints": c, "table_size": t,
                   "substeps": scene.substeps, "constraint_iters": iters,
                   "hose_per_frame": scene.hose["per_frame"] if scene.hose else 0},
        "hbm_model_bytes_per_substep": int(b_sub),
        "hbm_roofline_frac_model": round(model_frac, 4),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "serving": serving,
        "passes": passes,
        "pair_slots": inf_end.pair_slots,
        "pair_overflow": inf_end.pair_overflow,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
