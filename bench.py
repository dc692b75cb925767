"""Benchmark of the b2World::Step hot path on one MI355X.

Headline workload: the large random pile.
- 20,000 dynamic shapes: 60% convex polygons with 3-8 vertices, 40% circles.
- A static 200 x 200 m box of 4 edges.
- dt 1/60, 8 velocity / 3 position iterations, continuous physics and sleep
  off: the configuration the body-steps/s metric is quoted on.
- 500 steps: 20 warm-up steps through the drop-in b2World API, then 480
  timed steps through the b2g C ABI, launched kernel by kernel from the host
  with one colour read-back per step (--launch graph replays a captured step
  graph instead), with the world resident in HBM.
- `value` = dynamic bodies x timed steps / wall time.

Further sections of the same JSON line:
- `gas_chains`: 100k circles + 500 revolute chains, one world.
- `batched`: 4096 worlds, about 1.2M bodies, in one launch per stage,
  against a plain per-world loop; also the whole step's algorithmic HBM
  bytes over its time (`roofline_step`) and the reference CPU Step on a
  sample of the worlds, on one core and on all host cores.
- `serving`: Poisson requests (20k/s, 1-8 substeps at dt 1/60 or 1/240)
  against a resident pool of 10,000 worlds (16-4096 bodies, log-uniform)
  under a 64 GB memory budget; p50/p99 latency beside throughput.

roofline: HIP events around each run of consecutive launches of the
dominant kernel (one sweep's per-colour launches), over a further measured
segment, give its average launch duration. Algorithmic bytes
per launch are the per-item payloads of the reference's data layout (counted
once per pass) times the items the launch processed.

cpu_baseline: the reference library itself (oracle/_ref, compiled from the
reference sources) steps the same scene over the same steps on one host core.
The CPU restatement stands in where the reference library is not present.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

HBM_PEAK_GBS = 8000.0

# Per-item algorithmic bytes (reference FP32 payload sizes: manifold 64 B,
# velocity constraint 168 B, position constraint 88 B, body position or
# velocity 12 B, transform 16 B, AABB 16 B).
#   narrowphase (per contact): 2 transforms + 2 geometries + old manifold R + new W
#   contact velocity sweep (per solved contact): VC R + impulses W + 2 bodies' velocity R+W
#   contact position sweep (per solved contact): PC R + 2 bodies' position R+W
#   pair write (per shape): AABB R + key/index W (+ 8 B per reported pair)
def item_bytes(kernel: str, geo_bytes: float) -> float:
    return {
        "narrowphase": 2 * 16 + 2 * geo_bytes + 64 + 64,
        "contact_velocity": 168 + 16 + 2 * 2 * 12,
        "contact_position": 88 + 2 * 2 * 12,
        "pairs_write": 16 + 8,
    }[kernel]


def step_bytes(B: float, S: float, C: float, T: float, P: float, J: float, geo: float, V: int = 8, Q: int = 3):
    """Algorithmic bytes of one whole step, each object counted once per pass
    at the reference's FP32 payload sizes (the per-pass terms of the kernels
    above; Q = position iterations run, at most 3): B dynamic bodies, S shapes,
    C contacts, T touching contacts, P pairs, J joints, geo mean geometry."""
    return (C * (2 * 16 + 2 * geo + 64 + 64)              # narrowphase
            + B * (24 + 32 + 24)                           # integrate velocities
            + T * (64 + 2 * (24 + 16) + 168 + 88)          # contact init
            + T * (168 + 2 * 2 * 12)                       # warm start
            + V * (T * (168 + 16 + 2 * 2 * 12) + J * (112 + 2 * 2 * 12))
            + T * 16                                       # store impulses
            + B * (2 * 12 + 12)                            # integrate positions
            + Q * (T * (88 + 2 * 2 * 12) + J * (112 + 2 * 2 * 12))
            + B * (24 + 16)                                # write back state + transform
            + S * (16 + geo + 16)                          # AABB refresh
            + S * (16 + 8) + P * 8)                        # broadphase


def measured_traffic_ratio(kernel: str, workload: str):
    """HBM bytes / algorithmic bytes per launch of `kernel`, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE runs scaled by a plain
    copy kernel), or None when that kernel/workload was not measured."""
    path = os.path.join(REPO, "profiles", f"pmc_traffic_{workload}.md")
    if kernel != "contact_velocity" or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_ratio:"):
            return float(line.split(":")[1])
    return None


def workloads(pkg):
    S = pkg.scenes
    return {
        "pile20k": dict(make=lambda: S.random_pile(20000, seed=0xC2), desc="random pile, 20000 shapes "
                        "(60% polygons 3-8 verts r U[0.25,1], 40% circles r U[0.25,0.75]) in a 200x200 m "
                        "edge box, seed 0xC2", steps_total=500),
        "pyramid20": dict(make=lambda: S.pyramid(20), desc="pyramid of 20 rows (210 boxes) on an edge ground",
                          steps_total=1000),
        "gas100k_chains500": dict(make=lambda: S.gas_and_chains(), desc="100000 circles + 500 chains of 30 "
                                  "revolute-jointed boxes, zero gravity, 1000x1000 m box", steps_total=300),
    }


def mean_geometry_bytes(scene) -> float:
    import numpy as np
    kinds = np.asarray(scene.fix_kind)
    geo = np.asarray(scene.fix_geom)
    b = np.where(kinds == 0, 12.0, np.where(kinds == 1, 40.0, 0.0))
    poly = kinds == 2
    b[poly] = 16.0 * geo[poly, 0] + 16.0
    b[kinds == 3] = 16.0 * 4 + 16.0
    dyn_shapes = np.asarray(scene.body_type)[np.asarray(scene.fix_body)] == 2
    return float(b[dyn_shapes].mean()) if dyn_shapes.any() else float(b.mean())


def cpu_baseline(scene, warmup: int, steps: int):
    """The reference Step on one host core over the same simulation steps the
    GPU timed (warm-up steps untimed, as for the GPU)."""
    import numpy as np
    import oracle
    if oracle.Ref.available():
        w = oracle.RefWorld(scene)
        kind = "reference"
    else:
        w = oracle.RestatedWorld(scene)
        kind = "port"
    if warmup:
        w.step(warmup)
    t0 = time.perf_counter()
    if kind == "reference":
        ms = w.step(steps, timed=True)
    else:
        w.step(steps)
        ms = None
    el = time.perf_counter() - t0
    ndyn = scene.dynamic_count
    p99 = float(np.percentile(ms, 99)) if ms is not None else None
    return {"value": ndyn * steps / el, "unit": "body-steps/s", "cores": 1, "kind": kind,
            "sample": f"steps {warmup + 1}-{warmup + steps} of the same scene (the GPU's timed steps), "
                      f"single thread, {el:.1f} s wall, {el / steps * 1e3:.2f} ms/step mean"
                      + (f", p99 {p99:.2f} ms" if p99 is not None else "")}


def batched_section(pkg, worlds: int, warmup: int, steps: int, sample_every: int, cpu: bool, mode: int):
    """Batched pattern: `worlds` independent worlds (lognormal sizes, median
    200 bodies, pyramids and random piles) stepped by one set of launches per
    stage. Compared with a plain loop of single-world GPU steps (one world per
    batch, stepped one after another with a host sync per step, as a caller
    looping over b2World::Step would) on a sample of the same worlds, and with
    the reference CPU Step on the same sample. The warm-up is at least 12
    steps: the per-world solver's workgroup autotuner takes the first 10
    (eager steps, whatever the launch mode)."""
    import numpy as np
    S, eng = pkg.scenes, pkg.engine
    warmup = max(warmup, 12)
    sizes = S.batched_sizes(worlds)
    t0 = time.perf_counter()
    scenes = [S.batch_world(i, int(sizes[i])) for i in range(worlds)]
    batch = eng.Batch.from_scenes(scenes)
    build_s = time.perf_counter() - t0

    def timed(b, m):
        b.advance(max(warmup - 2, 1))
        for _ in range(2):  # capture the step graph of both contact-buffer parities
            b.step(8, 3, m & ~eng.STEP_ASYNC)
        b.wait()
        t = time.perf_counter()
        for _ in range(steps):
            b.step(8, 3, m)
        b.wait()
        return time.perf_counter() - t

    el = timed(batch, mode)
    counts = batch.counts()
    solver = {"path": "per-world workgroups" if batch.get_option(eng.OPT_LAST_SOLVER) == 2 else "colour sweeps",
              "tuned_workgroup": batch.get_option(eng.OPT_TUNED_WORLD_BLOCK),
              "tuned_lds_bodies": bool(batch.get_option(eng.OPT_TUNED_WORLD_LDS))}
    batch.close()
    # the same steps (a second batch of the same worlds, advanced alike) in
    # the other launch mode: replays of the captured step graph, or eager
    other = eng.STEP_ASYNC | (0 if mode & eng.STEP_GRAPH else eng.STEP_GRAPH)
    batch = eng.Batch.from_scenes(scenes)
    el_other = timed(batch, other)
    batch.close()
    nbodies = int(sum(s.dynamic_count for s in scenes))
    sample = list(range(0, worlds, sample_every))
    # plain per-world loop on the sample
    loop_steps = max(1, min(steps, 20))
    loop_time = 0.0
    for i in sample:
        b1 = eng.Batch.from_scenes([scenes[i]]).advance(max(warmup, 1))
        b1.wait()
        t1 = time.perf_counter()
        b1.advance(loop_steps)
        b1.wait()
        loop_time += time.perf_counter() - t1
        b1.close()
    loop_rate = len(sample) * loop_steps / loop_time
    out = {
        "metric": "batched world-steps/s",
        "value": worlds * steps / el,
        "unit": "world-steps/s",
        "body_steps_per_s": nbodies * steps / el,
        "ms_per_step": el / steps * 1e3,
        "worlds": worlds, "dynamic_bodies": nbodies, "steps": steps, "warmup": warmup,
        "contacts": counts.contacts, "touching": counts.touching, "colours": counts.colours,
        "solver": solver,
        "build_s": build_s,
        "per_world_loop": {"value": loop_rate, "unit": "world-steps/s",
                           "sample": f"{len(sample)} worlds (every {sample_every}th), {loop_steps} steps each, "
                                     f"one batch per world, host sync per step"},
        "speedup_vs_per_world_loop": (worlds * steps / el) / loop_rate,
        "launch": "hipGraph replay of the whole step" if mode & eng.STEP_GRAPH else "eager",
        "other_launch": {"launch": "eager" if mode & eng.STEP_GRAPH else "hipGraph replay of the whole step",
                         "ms_per_step": el_other / steps * 1e3, "value": worlds * steps / el_other,
                         "unit": "world-steps/s", "steps": steps,
                         "note": "a second batch of the same worlds, the same step range"},
    }
    # the whole step against HBM bandwidth: the genuinely HBM-sized case
    geo = float(np.mean([mean_geometry_bytes(sc) for sc in scenes]))
    nshapes = sum(sc.fixture_count for sc in scenes)
    sb = step_bytes(nbodies, nshapes, counts.contacts, counts.touching, counts.contacts, 0, geo)
    achieved = sb / (el / steps) / 1e9
    out["roofline_step"] = {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                            "frac": achieved / 8000.0, "bytes_per_step": sb,
                            "note": "whole-step algorithmic bytes (every pass of the step, position iterations "
                                    "counted as 3) over the measured step time"}
    if cpu:
        import oracle
        # the CPU sample: every 16th world, 200 steps each (a few seconds on one core)
        cpu_sample = list(range(0, worlds, 16))
        ref_steps = 200
        t_cpu = 0.0
        kind = "reference" if oracle.Ref.available() else "port"
        for i in cpu_sample:
            w = oracle.RefWorld(scenes[i]) if kind == "reference" else oracle.RestatedWorld(scenes[i])
            w.step(max(warmup, 1))
            t1 = time.perf_counter()
            w.step(ref_steps)
            t_cpu += time.perf_counter() - t1
        out["cpu_baseline"] = {"value": len(cpu_sample) * ref_steps / t_cpu, "unit": "world-steps/s", "cores": 1,
                               "kind": kind, "sample": f"{len(cpu_sample)} of the worlds (every 16th), "
                                                       f"{ref_steps} steps each, one thread, {t_cpu:.1f} s"}
        # the same worlds on all host cores, one world per task (built serially;
        # each task steps its world in one library call, outside the GIL)
        from concurrent.futures import ThreadPoolExecutor
        cores = max(1, min(16, os.cpu_count() or 1))
        mt = [oracle.RefWorld(scenes[i]) if kind == "reference" else oracle.RestatedWorld(scenes[i])
              for i in cpu_sample]
        for w in mt:
            w.step(max(warmup, 1))
        t1 = time.perf_counter()
        with ThreadPoolExecutor(cores) as ex:
            list(ex.map(lambda w: w.step(ref_steps), mt))
        t_mt = time.perf_counter() - t1
        out["cpu_baseline_all_cores"] = {"value": len(cpu_sample) * ref_steps / t_mt, "unit": "world-steps/s",
                                         "cores": cores, "kind": kind,
                                         "sample": f"the same {len(cpu_sample)} worlds, {ref_steps} steps each, "
                                                   f"{cores} threads, {t_mt:.2f} s"}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        out["speedup_vs_cpu_all_cores"] = out["value"] / out["cpu_baseline_all_cores"]["value"]
    return out


def single_world_section(pkg, name: str, warmup: int, steps: int, mode: int, cpu_steps: int):
    """Another single-world workload (default: 100k circles + 500 revolute
    chains of 30 links, zero gravity): GPU body-steps/s over `steps` steps
    after `warmup`, and the reference CPU Step over the same steps."""
    eng = pkg.engine
    wl = workloads(pkg)[name]
    scene = wl["make"]()
    world = eng.World(scene)
    world.step(max(warmup - 2, 1))
    batch = world.batch()
    import numpy as np
    dt = np.float32(1.0 / 60.0)
    ratio = np.float32(np.float32(1.0) / dt) * dt
    batch.set_world_params(0, [eng.WorldParams((scene.gravity[0], scene.gravity[1]), dt, ratio, 1, 1)])
    for _ in range(2):
        batch.step(8, 3, mode & ~eng.STEP_ASYNC)
    batch.wait()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.step(8, 3, mode)
    batch.wait()
    el = time.perf_counter() - t0
    c = batch.counts()
    out = {"workload": name, "description": wl["desc"], "value": scene.dynamic_count * steps / el,
           "unit": "body-steps/s", "ms_per_step": el / steps * 1e3, "steps": steps, "warmup": warmup,
           "dynamic_bodies": scene.dynamic_count, "joints": scene.joint_count, "contacts": c.contacts,
           "touching": c.touching, "colours": c.colours}
    if cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(wl["make"](), warmup, cpu_steps)
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def serving_section(pkg, worlds: int, rate: float, duration: float, budget_gb: float, graph: bool):
    """Serving pattern: Poisson world-step requests (1-8 substeps at dt 1/60 or
    1/240) against a resident pool of worlds with log-uniform sizes 16-4096,
    run by the native continuous-batching loop (b2s_run) on the wall clock.
    Latency = completion - arrival; throughput = body-steps completed / wall."""
    import numpy as np
    S, eng = pkg.scenes, pkg.engine
    sizes = S.serving_sizes(worlds, seed=0xC5)
    t0 = time.perf_counter()
    scenes = [S.batch_world(50000 + i, int(sizes[i]), seed_base=0xC5) for i in range(worlds)]
    pool = eng.Batch.from_scenes(scenes).advance(1)
    pool.wait()
    build_s = time.perf_counter() - t0
    wbytes = np.array([eng.world_device_bytes(s) for s in scenes], np.int64)
    trace = S.serving_trace(worlds, rate=rate, duration=duration, seed=0xC5)
    res, st, _ = eng.serve(pool, trace, int(budget_gb * (1 << 30)), wbytes, graph=graph, realtime=True)
    lat = (res["completed"] - res["arrival"]) * 1e3
    queue = (res["admitted"] - res["arrival"]) * 1e3
    out = {
        "metric": "serving p99 request latency", "value": float(np.percentile(lat, 99)), "unit": "ms",
        "higher_is_better": False,
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "p50_queue_ms": float(np.percentile(queue, 50)), "p99_queue_ms": float(np.percentile(queue, 99)),
        "body_steps_per_s": st.body_steps / st.wall_s, "world_steps_per_s": st.world_steps / st.wall_s,
        "requests": int(len(trace)), "completed": int(st.completed), "offered_rate_per_s": rate,
        "duration_s": duration, "wall_s": st.wall_s, "ticks": int(st.ticks),
        "mean_in_flight": st.mean_in_flight, "max_in_flight": int(st.max_in_flight),
        # device time of the ticks (HIP events around each step) over the wall
        # time, and the host's time inside the step calls
        "device_busy_frac": st.device_s / st.wall_s if st.device_s > 0 else None,
        "host_step_frac": st.busy_s / st.wall_s,
        "pool": {"worlds": worlds, "bodies": int(sum(s.body_count for s in scenes)),
                 "sizes": "log-uniform 16-4096", "memory_budget_gb": budget_gb, "build_s": build_s},
        "launch": "hipGraph replay" if graph else "eager",
    }
    pool.close()
    return out


def copy_probe(mib: int, reps: int) -> None:
    """A plain device-to-device copy of `mib` MiB, `reps` times: the reference
    kernel for scaling the FETCH_SIZE / WRITE_SIZE counters of a --pmc run
    (each dispatch reads and writes exactly mib MiB)."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"copy_probe_mib": mib, "reps": reps, "GB_per_s": 2 * mib * (1 << 20) * reps / el / 1e9}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=480, help="timed steps (with the warm-up: the 500-step run)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="pile20k")
    ap.add_argument("--kernel", default="contact_velocity",
                    choices=["narrowphase", "contact_velocity", "contact_position", "pairs_write"])
    ap.add_argument("--kernel-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=-1, help="-1: the GPU's timed steps (bounded)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--eager", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--launch", choices=["eager", "graph"], default="eager",
                    help="timed steps: kernels launched from the host with one colour read-back per step "
                         "(eager, default) or a replayed hipGraph of the whole step (graph)")
    ap.add_argument("--solver", choices=["auto", "colours", "cells"], default="auto",
                    help="contact solver of the headline world (auto: the colour sweeps for one world)")
    ap.add_argument("--batch-worlds", type=int, default=4096, help="0 skips the batched section")
    ap.add_argument("--batch-steps", type=int, default=50)
    ap.add_argument("--batch-sample", type=int, default=64, help="per-world loop / CPU sample stride")
    ap.add_argument("--serve-worlds", type=int, default=10000, help="0 skips the serving section")
    ap.add_argument("--gas-steps", type=int, default=100, help="0 skips the gas + chains section")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration) and exit")
    ap.add_argument("--serve-rate", type=float, default=20000.0)
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    args = ap.parse_args()
    if args.eager:
        args.launch = "eager"
    args.eager = args.launch == "eager"
    if args.copy_probe:
        copy_probe(args.copy_probe, 10)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: worlds shard by world index across GPUs with no "
                                   "exchange, but no multi-GPU run is part of it; use --gpus 1"}))
        sys.exit(2)

    from util import load_pkg
    import numpy as np
    pkg = load_pkg()
    eng = pkg.engine
    wl = workloads(pkg)[args.workload]
    scene = wl["make"]()
    ndyn = scene.dynamic_count

    mode = eng.STEP_ASYNC | (0 if args.eager else eng.STEP_GRAPH)
    world = eng.World(scene)
    world.step(max(args.warmup - 2, 1))
    batch = world.batch()
    dt = np.float32(1.0 / 60.0)
    ratio = np.float32(np.float32(1.0) / dt) * dt
    batch.set_world_params(0, [eng.WorldParams((scene.gravity[0], scene.gravity[1]), dt, ratio, 1, 1)])
    batch.set_option(eng.OPT_SOLVER, {"auto": eng.SOLVER_AUTO, "colours": eng.SOLVER_COLOURS,
                                      "cells": eng.SOLVER_CELLS}[args.solver])
    for _ in range(2):  # last warm-up steps: capture the step graph of both contact-buffer parities
        batch.step(8, 3, mode & ~eng.STEP_ASYNC)
    batch.wait()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.step(8, 3, mode)
    batch.wait()
    elapsed = time.perf_counter() - t0
    counts = batch.counts()
    solver_path = {1: "colour sweeps", 2: "per-world workgroups", 3: "cells"}.get(
        batch.get_option(eng.OPT_LAST_SOLVER), "unknown")

    kid = {"narrowphase": eng.KERNEL_NARROWPHASE, "contact_velocity": eng.KERNEL_CONTACT_VELOCITY,
           "contact_position": eng.KERNEL_CONTACT_POSITION, "pairs_write": eng.KERNEL_PAIRS_WRITE}[args.kernel]
    batch.time_kernel(kid)
    for _ in range(args.kernel_steps):
        batch.step(8, 3, eng.STEP_ASYNC)
    batch.wait()
    k_ms, k_launches, k_items = batch.kernel_time()
    batch.time_kernel(-1)
    geo = mean_geometry_bytes(scene)
    per_item = item_bytes(args.kernel, geo)
    avg_ms = k_ms / max(k_launches, 1)
    bytes_per_launch = per_item * k_items / max(k_launches, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0

    cpu = None
    if not args.no_cpu:
        cpu_steps = args.cpu_steps
        if cpu_steps < 0:
            cpu_steps = min(args.steps, 480 if args.workload == "pile20k" else 1000 if args.workload == "pyramid20"
                            else 40)
        cpu = cpu_baseline(wl["make"](), args.warmup, cpu_steps)

    batched = None
    if args.batch_worlds > 0:
        batched = batched_section(pkg, args.batch_worlds, args.warmup, args.batch_steps, args.batch_sample,
                                  not args.no_cpu, mode)

    gas = None
    if args.gas_steps > 0:
        gas = single_world_section(pkg, "gas100k_chains500", args.warmup, args.gas_steps, mode,
                                   0 if args.no_cpu else min(args.gas_steps, 40))

    serving = None
    if args.serve_worlds > 0:
        serving = serving_section(pkg, args.serve_worlds, args.serve_rate, args.serve_seconds,
                                  args.serve_budget_gb, not args.eager)

    ratio = measured_traffic_ratio(args.kernel, args.workload)
    value = ndyn * args.steps / elapsed
    line = {
        "metric": "body-steps/s on 1 GPU vs host CPU (20k pile)",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": args.workload, "description": wl["desc"], "dynamic_bodies": ndyn,
                   "shapes": scene.fixture_count, "joints": scene.joint_count, "dt": "1/60",
                   "velocity_iterations": 8, "position_iterations": 3,
                   "contacts": counts.contacts, "touching": counts.touching, "colours": counts.colours,
                   "steps_timed": f"{args.warmup + 1}-{args.warmup + args.steps}",
                   "launch": "eager" if args.eager else "hipGraph replay of the whole step",
                   "graph_steps": counts.graph_steps, "serial_constraints": counts.serial,
                   "solver": solver_path},
        "roofline": {"bound": "hbm", "kernel": args.kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (ratio * bytes_per_launch) if ratio else None,
                     "traffic_source": (f"profiles/pmc_traffic_{args.workload}.md: measured HBM/algorithmic "
                                        f"ratio {ratio:.2f} x this run's algorithmic bytes per launch")
                     if ratio else None,
                     "avg_launch_ms": avg_ms, "launches": k_launches,
                     "bytes_per_launch": bytes_per_launch, "bytes_per_item": per_item},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
        "gas_chains": gas,
        "batched": batched,
        "serving": serving,
        "host": platform.node(),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
