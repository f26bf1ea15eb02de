"""Benchmark: body-steps/s of World.Step on one MI355X for a 131,072-body heap, and iterations to a
1e-3 velocity residual with the learned vs the previous-frame (plain) warm start.

Workload: `scenes.heap(131072)` — dynamic spheres/boxes/capsules dropped into a static container,
dt = 1/60, 4 substeps, 10 solver iterations, 4 relaxation iterations (the 32,768-body heap
configuration scaled to 131,072 bodies, the body count the metric is quoted at).  W warmup steps let
the heap form contacts; then K steps are timed with inputs resident in HBM.

Besides the JSON line's contract fields it reports:
  roofline     — the solver kernel (k_solve_flow, the persistent dataflow solver; k_solve_tiles, one
                 workgroup per Morton range of bodies; or k_solve_worlds, one workgroup per world,
                 whichever the autotuner picked), whose duration is measured with HIP events around
                 its launch on the library's stream (jh_stats.ms_solve_kernel); algorithmic bytes per
                 launch from the per-step counts (see solver_bytes()); traffic from the committed
                 counter summary when this is the profiled run (see measured_traffic()).
  cpu_baseline — the oracle (CPU restatement, 1 thread) timed on a bounded sample: a few steps of the
                 same heap continued from the GPU state at the start of the timed region.
  learned_warm_start — residual probes of the settled heap (see learned_warm_start_leg).

Other workloads (--workload): pyramid, batched, full (the other configurations), joints (4,096 jointed
groups with constraints of every kind, solved by the barrier solver k_solve) and bullets (32,768
projectiles with speculative contacts); their roofline object names the solver kernel that ran.
"""
from __future__ import annotations

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


class _Roctx:
    """roctx range around the timed steps, so a kernel-trace profile of this command can be cut to
    them (rocpd2summary --start-marker/--end-marker timed_steps).  No-op without the library."""

    def __init__(self):
        import ctypes
        self.lib = None
        for name in ("librocprofiler-sdk-roctx.so.1", "libroctx64.so.4"):
            try:
                self.lib = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if self.lib is not None:
            self.lib.roctxRangePushA.argtypes = [ctypes.c_char_p]

    def push(self, name: str):
        if self.lib is not None:
            self.lib.roctxRangePushA(name.encode())

    def pop(self):
        if self.lib is not None:
            self.lib.roctxRangePop()


def solver_bytes(stats: dict, n_dyn: int, n_moving: int, substeps: int, iters: int, relax: int,
                 n_cons: int = 0) -> float:
    """Algorithmic HBM bytes of one solver launch (one step, all substeps).
    Per arbiter per pass (prepare, each iteration, each relaxation): 152 B of body traffic
    (2 x (v, w read + write 48 B + inverse mass/inertia 28 B)); per contact point per pass 116 B
    (row data 92 B read, accumulated impulse 12 B read + 12 B write).  Per constraint per pass
    (prepare and each iteration; relaxation skips constraints): 544 B (record 112 B read, state 32 B
    write, work record 192 B, 2 x 104 B of body velocities, inverse mass/inertia, position and
    orientation).  Per substep: IntegrateForces 72 B per dynamic body, IntegrateVelocities 80 B per
    non-static body."""
    m, k = stats["arbiters_solved"], stats["contacts"]
    passes = 1 + iters + relax
    per_substep = passes * (152.0 * m + 116.0 * k) + (1 + iters) * 544.0 * n_cons + 72.0 * n_dyn + 80.0 * n_moving
    return substeps * per_substep


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_solve_flow_heap131072_traffic.json")


def measured_traffic(kernel: str, args) -> float | None:
    """HBM bytes per launch of the solver kernel from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by what it reads for a device-to-device copy of known
    size).  That summary profiles the default run (131,072-body heap, 240 warmup steps, autotune,
    60 timed steps, dataflow solver); any other run reports None."""
    if (kernel != "k_solve_flow" or args.workload != "heap" or args.bodies != 131072 or args.warmup != 240
            or args.steps != 60 or not os.path.exists(TRAFFIC_PROFILE)):
        return None
    with open(TRAFFIC_PROFILE) as f:
        prof = json.load(f)
    return float(prof["traffic_bytes_per_launch"]) if prof.get("kernel") == kernel else None


def learned_warm_start_leg(gw, dt: float, cfg, settle_to: int, probe_iterations: int, timed: int) -> dict:
    """iters to 1e-3 residual, learned vs previous-frame warm start, on the benchmark heap once it
    has settled (step `settle_to`): one probe step runs the first substep from identical states with
    each warm start (and a cold start, the normaliser of R_k).  Then, at equal accuracy (the learned
    warm start with the fewest iterations whose residual is at most the previous-frame one's at the
    configured count), `timed` steps of each are timed."""
    from jitterphysics2_amd.lws import (PROBE_COLD, PROBE_LEARNED, PROBE_PREVIOUS, LearnedWarmStart, iterations_to,
                                        relative_residuals)
    from jitterphysics2_amd.lws_train import DEFAULT_MODEL

    if not os.path.exists(DEFAULT_MODEL):
        return {"error": f"no trained model at {DEFAULT_MODEL} (python -m jitterphysics2_amd.lws_train)"}
    model = LearnedWarmStart.load(DEFAULT_MODEL)
    gw.set_warm_start_model(model)
    while gw.stats()["steps"] < settle_to:
        gw.step(dt)
    r = gw.step_probe(dt, probe_iterations, PROBE_PREVIOUS | PROBE_LEARNED | PROBE_COLD)
    R = relative_residuals(r["residual"])
    it = int(cfg["solver_iterations"][0])
    out = {"state": f"benchmark heap after {settle_to} steps (settled)", "contact_points": r["points"],
           "probe_iterations": probe_iterations, "tolerance": 1e-3}
    for m, name in ((0, "previous"), (1, "learned"), (2, "cold")):
        out[name] = {"iters_to_1e-3": iterations_to(R[m], 1e-3),
                     "R": {str(k): float(R[m, k]) for k in (0, 4, 8, 16, 32, 64, probe_iterations)}}
    k_learned = iterations_to(R[1], float(R[0, it]))
    out["equal_accuracy"] = {"previous_iterations": it, "R_previous": float(R[0, it]), "learned_iterations": k_learned}
    if k_learned is not None and k_learned >= 1 and timed > 0:
        rates = {}
        for name, ws, iters in (("previous", 0, it), ("learned", 1, k_learned)):
            c = cfg.copy()
            c["warm_start"] = ws
            c["solver_iterations"] = iters
            gw.set_config(c)
            n_dyn = gw.stats()["bodies"]
            t0 = time.perf_counter()
            for _ in range(timed):
                gw.step(dt)
            rates[name] = {"iterations": iters, "body_steps_per_s": n_dyn * timed / (time.perf_counter() - t0)}
        gw.set_config(cfg)
        out["equal_accuracy"]["throughput"] = rates
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (heap default 60; pyramid 600)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps first (heap default 240; pyramid 0)")
    ap.add_argument("--bodies", type=int, default=131072)
    ap.add_argument("--workload", choices=("heap", "batched", "full", "pyramid", "joints", "bullets"), default="heap",
                    help="heap: one heap of --bodies bodies (default 131,072, the metric's configuration; 32,768 is "
                         "the 32k-heap configuration); batched: 4,096 independent worlds of 16..2,048 bodies "
                         "(log-uniform) in one world batch; full: 262,144 bodies incl. 32-point hulls over a "
                         "512x512-cell triangle heightfield split into 64 sub-worlds, TriangleEdgeCollisionFilter "
                         "on; pyramid: the 20-row pyramid of 210 unit boxes, 600 steps, CPU baseline over all of "
                         "them; joints: --groups jointed groups (hinge chains, ropes, welds, distance limits, "
                         "limbs, motors, sliders, universal joints, servos) over a floor; bullets: --bodies "
                         "projectiles at 15..60 m/s with speculative contacts against thin slabs")
    ap.add_argument("--groups", type=int, default=4096, help="joints workload: jointed groups")
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=None, help="CPU baseline sample (default 3 steps; joints 20)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lws", action="store_true", help="skip the learned-warm-start residual leg")
    ap.add_argument("--lws-settle", type=int, default=700)
    ap.add_argument("--lws-iterations", type=int, default=256)
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 600 if args.workload == "pyramid" else 60
    if args.warmup is None:  # bullets: time the impacts, not the pile they leave
        args.warmup = {"pyramid": 0, "bullets": 10}.get(args.workload, 240)
    if args.cpu_steps is None:
        args.cpu_steps = {"joints": 20}.get(args.workload, 3)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described in DESIGN.md",
                          "n_gpus": args.gpus}))
        return 2

    from jitterphysics2_amd import scenes
    from jitterphysics2_amd.gpu import GpuWorld

    t0 = time.time()
    if args.workload == "batched":
        bodies, shapes, verts, cfg, dt = scenes.batched_worlds(args.worlds)
        args.no_lws = True  # the residual leg is defined on the heap
    elif args.workload == "full":
        bodies, shapes, verts, cfg, dt = scenes.full_form()
        args.no_lws = True
    elif args.workload == "pyramid":
        bodies, shapes, verts, cfg, dt = scenes.pyramid(rows=20)
        args.no_lws = True
        args.cpu_steps = args.steps  # the whole configuration is CPU-sized
        args.no_autotune = True  # the configuration's 600 steps start from the initial state
    elif args.workload == "joints":
        bodies, shapes, verts, cfg, dt, cons = scenes.joints(args.groups)
        args.no_lws = True
    elif args.workload == "bullets":
        n = 32768 if args.bodies == 131072 else args.bodies
        args.bodies = n
        bodies, shapes, verts, cfg, dt = scenes.bullets(n, arena=30.0 * math.sqrt(n / 512.0))
        args.no_lws = True
    else:
        bodies, shapes, verts, cfg, dt = scenes.heap(args.bodies)
    if args.workload != "joints":
        cons = None
    print(f"# scene built: {len(bodies)} bodies, {len(shapes)} shapes in {time.time() - t0:.1f}s", flush=True)
    n_dyn = int(np.sum(bodies["motion_type"] == 0))
    n_moving = int(np.sum(bodies["motion_type"] != 2))
    substeps = int(cfg["substep_count"][0])
    iters = int(cfg["solver_iterations"][0])
    relax = int(cfg["relaxation_iterations"][0])

    gw = GpuWorld(bodies, shapes, verts, cfg)
    if cons is not None:
        gw.set_constraints(cons)
    t0 = time.time()
    for i in range(args.warmup):
        gw.step(dt)
        if (i + 1) % 60 == 0:
            s = gw.stats()
            print(f"# warmup {i + 1}/{args.warmup}: pairs {s['pairs']} arbiters {s['arbiters']} "
                  f"colours {s['colours']} ms/step {s['ms_step']:.2f}", flush=True)
    print(f"# warmup done in {time.time() - t0:.1f}s", flush=True)
    tuning = None
    if not args.no_autotune:
        tuning = gw.autotune(dt, rounds=2)
        print(f"# autotune: solver {tuning['solver']} x {tuning['solve_blocks']} ({tuning['ms_solve_best']:.2f} ms vs "
              f"{tuning['ms_solve_default']:.2f} default), colour blocks {tuning['colour_blocks']} "
              f"({tuning['ms_colour_best']:.2f} ms vs {tuning['ms_colour_default']:.2f})", flush=True)

    start_bodies = gw.bodies()
    start_arbiters = gw.arbiters()
    start_cons = gw.constraints() if cons is not None else None
    con_order = gw.constraint_order()[0] if cons is not None else None
    n_cons = 0 if cons is None else int(np.sum(cons["enabled"] != 0))

    solve_ms, step_ms, bytes_per_launch, contacts = [], [], [], []
    stage_keys = ("ms_broadphase", "ms_narrowphase", "ms_epa", "ms_contact_cache", "ms_colouring", "ms_solve", "ms_update")
    stage_sum = dict.fromkeys(stage_keys, 0.0)
    marker = _Roctx()
    marker.push("timed_steps")
    t_start = time.perf_counter()
    for _ in range(args.steps):
        gw.step(dt)  # returns after the step finished on the GPU
        s = gw.stats()
        solve_ms.append(s["ms_solve_kernel"])
        for key in stage_keys:
            stage_sum[key] += s[key]
        step_ms.append(s["ms_step"])
        contacts.append(s["contacts"])
        bytes_per_launch.append(solver_bytes(s, n_dyn, n_moving, substeps, iters, relax, n_cons))
    t_end = time.perf_counter()
    marker.pop()
    wall = t_end - t_start
    value = n_dyn * args.steps / wall
    ms_per_step = 1000.0 * wall / args.steps
    avg_solve_s = float(np.mean(solve_ms)) / 1000.0
    achieved = float(np.mean(bytes_per_launch)) / avg_solve_s / 1e9
    final = gw.stats()
    lws = None if args.no_lws else learned_warm_start_leg(gw, dt, cfg, max(args.lws_settle, args.warmup + args.steps),
                                                          args.lws_iterations, 20)
    gw.close()

    cpu = None
    if not args.no_cpu:
        from tests.orc import OracleWorld

        ow = OracleWorld(start_bodies, shapes, verts, cfg)
        ow.set_arbiters(start_arbiters)
        if start_cons is not None:
            ow.set_constraints(start_cons)
            ow.set_constraint_order(con_order)
        secs = ow.time_steps(dt, args.cpu_steps)
        cpu = {"value": n_dyn * args.cpu_steps / secs, "unit": "body-steps/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} steps of the same {n_dyn}-body scene continued from the GPU state at "
                         f"the start of the timed region (oracle FP32, key-order solve, 1 thread)"}

    kernel = {0: "k_solve", 1: "k_solve_flow", 2: "k_solve_worlds", 3: "k_solve_tiles",
              4: "k_solve_lanes"}[int(final["solver"])]
    traffic = measured_traffic(kernel, args)

    line = {
        "metric": "body-steps/sec at 100k bodies; iters to 1e-3 residual, learned vs plain warm start",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"heap of {n_dyn} dynamic spheres/boxes/capsules in a static container, dt=1/60, "
                                f"{substeps} substeps, {iters} iterations, {relax} relaxations (plain warm start)")
                   if args.workload == "heap" else
                   (f"{args.worlds} independent worlds, {n_dyn} dynamic boxes/spheres in total (16..2048 per "
                    f"world, log-uniform), dt=1/60, {substeps} substeps, {iters} iterations, {relax} relaxations")
                   if args.workload == "batched" else
                   (f"{n_dyn} dynamic spheres/boxes/capsules/32-point hulls over a 512x512-cell Perlin heightfield "
                    f"(524,288 triangles) in 64 sub-worlds, triangle edge filter, dt=1/60, {substeps} substeps, "
                    f"{iters} iterations, {relax} relaxations")
                   if args.workload == "full" else
                   (f"20-row pyramid of {n_dyn} unit boxes on a static ground box, dt=0.01, {substeps} substeps, "
                    f"{iters} iterations, {relax} relaxations")
                   if args.workload == "pyramid" else
                   (f"{args.groups} jointed groups, {n_dyn} dynamic bodies, {n_cons} constraints of 15 kinds over a "
                    f"static floor, dt=1/60, {substeps} substeps, {iters} iterations, {relax} relaxations")
                   if args.workload == "joints" else
                   (f"{n_dyn} projectiles (spheres/boxes/capsules/16-point hulls) at 15..60 m/s with speculative "
                    f"contacts against 0.2-thick static slabs, dt=1/60, {substeps} substeps, {iters} iterations, "
                    f"{relax} relaxations"),
                   "constraints": n_cons,
                   "autotune": tuning,
                   "bodies": n_dyn, "pairs": final["pairs"], "arbiters": final["arbiters"],
                   "contacts": final["contacts"], "colours": final["colours"], "max_degree": final["max_degree"],
                   "colliding_pairs": final["colliding_pairs"], "epa_pairs": final["epa_calls"],
                   "contacts_per_s": float(np.mean(contacts)) * args.steps / wall,
                   "stage_ms": {k: stage_sum[k] / args.steps for k in stage_keys}},  # mean over the timed steps
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "avg_launch_ms": 1000.0 * avg_solve_s,
                     "algorithmic_bytes_per_launch": float(np.mean(bytes_per_launch))},
        "cpu_baseline": cpu,
        "learned_warm_start": lws,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
