"""Benchmark: body-steps/s of World.Step on one MI355X for a 131,072-body heap, and iterations to a
1e-3 velocity residual with the learned vs the previous-frame (plain) warm start.

Workload: `scenes.heap(131072)` — dynamic spheres/boxes/capsules dropped into a static container,
dt = 1/60, 4 substeps, 10 solver iterations, 4 relaxation iterations (the 32,768-body heap
configuration scaled to 131,072 bodies, the body count the metric is quoted at).  The bodies start on a
grid up to 182 m high, so the heap first settles for --settle untimed steps (600 by default: the top
layers land after about 370), whatever --warmup is; then W warmup steps, the autotuner, and K timed
steps with inputs resident in HBM.

Besides the JSON line's contract fields it reports:
  roofline     — the solver kernel (k_solve_flow, the persistent dataflow solver; k_solve_tiles, one
                 workgroup per Morton range of bodies; or k_solve_worlds, one workgroup per world,
                 whichever the autotuner picked), whose duration is measured with HIP events around
                 its launch on the library's stream (jh_stats.ms_solve_kernel); algorithmic bytes per
                 launch from the per-step counts (see solver_bytes()); traffic from the committed
                 counter summary when this is the profiled run (see measured_traffic()).
  cpu_baseline — the oracle (CPU restatement) on the host's threads (OMP_NUM_THREADS), timed on a
                 bounded sample: a few steps of the same heap continued from the GPU state at the
                 start of the timed region (the narrowphase and per-body work in parallel, the
                 Gauss-Seidel solve sequential).  Its fp64_error: the FP32 error of one GPU step
                 (taken just before the timed region) against the FP64 oracle stepping the same state
                 in the GPU's colour order, over every dynamic body.
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
    (FETCH_SIZE + WRITE_SIZE over the timed steps only, each scaled by what it reads for a
    device-to-device copy of known size).  The summary records the workload it profiled (workload,
    bodies, settle, warmup and timed steps); a run of any other workload reports None."""
    if not os.path.exists(TRAFFIC_PROFILE):
        return None
    with open(TRAFFIC_PROFILE) as f:
        prof = json.load(f)
    run = prof.get("run", {})
    same = (prof.get("kernel") == kernel and run.get("workload") == args.workload and run.get("bodies") == args.bodies
            and run.get("settle") == args.settle and run.get("warmup") == args.warmup and run.get("steps") == args.steps)
    return float(prof["traffic_bytes_per_launch"]) if same else None


def calibrate_copy(mib: int) -> None:
    """One device-to-device copy of `mib` MiB (hipMemcpy, the runtime's copy kernel) before the scene:
    HBM traffic of known size, which scales the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    import ctypes
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
    hip.hipMemset.argtypes = [vp, ctypes.c_int, sz]
    hip.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
    hip.hipFree.argtypes = [vp]
    n = int(mib) << 20
    src, dst = vp(), vp()
    if hip.hipMalloc(ctypes.byref(src), n) or hip.hipMalloc(ctypes.byref(dst), n):
        raise RuntimeError("calibration copy: hipMalloc failed")
    if hip.hipMemset(src, 1, n) or hip.hipMemcpy(dst, src, n, 3) or hip.hipDeviceSynchronize():
        raise RuntimeError("calibration copy failed")
    hip.hipFree(src)
    hip.hipFree(dst)


def _csv_rows(path: str) -> list[dict]:
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _kname(r: dict) -> str:
    return r["Kernel_Name"].split("(")[0]


def trace_summary(trace_csv: str, steps: int, top: int = 16) -> str:
    """Per-kernel table of the timed steps of a `rocprofv3 --kernel-trace --output-format csv` run of
    this script with --no-lws --no-cpu: the window starts after the solver launch that precedes the
    last `steps` solver launches (settle, warmup and autotune steps fall before it)."""
    rows = sorted(_csv_rows(trace_csv), key=lambda r: int(r["Start_Timestamp"]))
    solves = [r for r in rows if _kname(r).startswith("jh::k_solve")]
    if len(solves) <= steps:
        raise SystemExit(f"{len(solves)} solver launches, need more than {steps}")
    cut = int(solves[-steps - 1]["End_Timestamp"])
    timed = [r for r in rows if int(r["Start_Timestamp"]) > cut]
    agg: dict = {}
    for r in timed:
        d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        wg = int(r.get("Workgroup_Size_X") or r.get("Workgroup_Size") or 1)
        grid = int(r.get("Grid_Size_X") or r.get("Grid_Size") or 0)
        a = agg.setdefault(_kname(r), [0, 0, grid // max(wg, 1), wg, r.get("VGPR_Count", ""),
                                        r.get("Accum_VGPR_Count", ""), r.get("LDS_Block_Size", "")])
        a[0] += 1
        a[1] += d
    busy = sum(a[1] for a in agg.values())
    span = int(timed[-1]["End_Timestamp"]) - int(timed[0]["Start_Timestamp"])
    out = [f"timed window: {len(timed)} dispatches, GPU busy {busy / 1e6:.2f} ms of {span / 1e6:.2f} ms wall "
           f"({steps} steps)",
           f"{'kernel':34s} {'calls':>6s} {'avg_us':>10s} {'per_step_ms':>12s} {'share':>7s} {'wgs':>6s} {'wg':>5s} "
           f"{'vgpr':>5s} {'agpr':>5s} {'lds':>6s}"]
    for n, a in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        out.append(f"{n:34s} {a[0]:6d} {a[1] / a[0] / 1e3:10.2f} {a[1] / steps / 1e6:12.3f} {100 * a[1] / busy:6.1f}% "
                   f"{a[2]:6d} {a[3]:5d} {a[4]:>5s} {a[5]:>5s} {a[6]:>6s}")
    return "\n".join(out)


def counter_summary(counter_csv: str, kernel: str, steps: int, copy_mib: int) -> dict:
    """Mean of each counter over the last `steps` launches of `kernel` (the timed steps), and the same
    counter for the calibration copy of `copy_mib` MiB (runtime copy kernels before the first jh::
    kernel): scale = bytes copied / counter bytes of the copy."""
    rows = sorted(_csv_rows(counter_csv), key=lambda r: int(r["Dispatch_Id"]))
    out: dict = {"kernel": kernel, "steps": steps, "copy_mib": copy_mib}
    first_jh = min(int(r["Dispatch_Id"]) for r in rows if _kname(r).startswith("jh::"))
    for name in sorted({r["Counter_Name"] for r in rows}):
        rk = [r for r in rows if r["Counter_Name"] == name and _kname(r).endswith(kernel)]
        per = {}
        for r in rk:
            per[int(r["Dispatch_Id"])] = per.get(int(r["Dispatch_Id"]), 0.0) + float(r["Counter_Value"])
        last = [per[d] for d in sorted(per)][-steps:]
        copy = sum(float(r["Counter_Value"]) for r in rows if r["Counter_Name"] == name and
                   int(r["Dispatch_Id"]) < first_jh and "copyBuffer" in r["Kernel_Name"])
        out[name] = {"kib_per_launch": float(np.mean(last)), "launches": len(last), "copy_kib": copy,
                     "scale": (copy_mib * 1024.0 / copy) if copy else None}
    return out


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
           "probe_iterations": probe_iterations, "tolerance": 1e-3,
           "tolerance_reached": 1e-2, "note": "iters_to_1e-2: a tolerance the warm-started solves reach within "
                                               "the probe (neither reaches 1e-3 on this heap)"}
    for m, name in ((0, "previous"), (1, "learned"), (2, "cold")):
        out[name] = {"iters_to_1e-3": iterations_to(R[m], 1e-3), "iters_to_1e-2": iterations_to(R[m], 1e-2),
                     "R": {str(k): float(R[m, k]) for k in (0, 4, 8, 16, 32, 64, probe_iterations)}}
    a, b = out["previous"]["iters_to_1e-2"], out["learned"]["iters_to_1e-2"]
    out["gain_iters_to_1e-2"] = (a / b) if a and b else None  # previous-frame iterations / learned iterations
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
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps before the autotuner and the timed "
                                                              "steps, after --settle (default 5; pyramid 0)")
    ap.add_argument("--settle", type=int, default=None, help="untimed steps that let the scene form its contacts "
                                                              "before everything else (heap default 600; others 0)")
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
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (scales HBM traffic counters of a profiled run)")
    ap.add_argument("--trace-summary", metavar="KERNEL_TRACE_CSV",
                    help="print the timed steps' kernel table of a rocprofv3 kernel trace of this script, then exit")
    ap.add_argument("--counter-summary", metavar="COUNTER_CSV",
                    help="print the timed steps' counters of the solver kernel (rocprofv3 --pmc), then exit")
    ap.add_argument("--kernel", default="k_solve_flow", help="kernel for --counter-summary")
    args = ap.parse_args()
    if args.trace_summary:
        print(trace_summary(args.trace_summary, args.steps or 60))
        return 0
    if args.counter_summary:
        print(json.dumps(counter_summary(args.counter_summary, args.kernel, args.steps or 60, args.calibrate_copy),
                         indent=1))
        return 0
    if args.steps is None:
        args.steps = 600 if args.workload == "pyramid" else 60
    if args.settle is None:  # the heap is timed once it has formed, not while it falls
        args.settle = 600 if args.workload == "heap" else 0
    if args.warmup is None:  # bullets: time the impacts, not the pile they leave
        args.warmup = {"pyramid": 0, "bullets": 10, "heap": 5}.get(args.workload, 240)
    if args.cpu_steps is None:
        args.cpu_steps = {"joints": 20}.get(args.workload, 3)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described in DESIGN.md",
                          "n_gpus": args.gpus}))
        return 2

    from jitterphysics2_amd import scenes
    from jitterphysics2_amd.gpu import GpuWorld

    if args.calibrate_copy > 0:
        calibrate_copy(args.calibrate_copy)
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
    import ctypes
    print(f"# native library {gw.lib._name} (jh_version at "
          f"{ctypes.cast(gw.lib.jh_version, ctypes.c_void_p).value:#x})", flush=True)
    if cons is not None:
        gw.set_constraints(cons)
    t0 = time.time()
    for i in range(args.settle + args.warmup):
        gw.step(dt)
        if (i + 1) % 100 == 0:
            s = gw.stats()
            print(f"# settle/warmup {i + 1}/{args.settle + args.warmup}: pairs {s['pairs']} arbiters {s['arbiters']} "
                  f"contacts {s['contacts']} colours {s['colours']} ms/step {s['ms_step']:.2f}", flush=True)
    print(f"# settle + warmup done in {time.time() - t0:.1f}s", flush=True)
    tuning = None
    if not args.no_autotune:
        tuning = gw.autotune(dt, rounds=2)
        print(f"# autotune: solver {tuning['solver']} x {tuning['solve_blocks']} ({tuning['ms_solve_best']:.2f} ms vs "
              f"{tuning['ms_solve_default']:.2f} default), colour blocks {tuning['colour_blocks']} "
              f"({tuning['ms_colour_best']:.2f} ms vs {tuning['ms_colour_default']:.2f}), traversal "
              f"{'wavefront, %d leaves per wave' % tuning['traversal'] if tuning['traversal'] else 'per-thread'} (broadphase {tuning['ms_broadphase_best']:.2f} ms "
              f"vs {tuning['ms_broadphase_default']:.2f})", flush=True)

    ref = None
    if not args.no_cpu:
        # one untimed step whose FP32 result the FP64 oracle checks (cpu_baseline.fp64_error)
        ref = {"bodies": gw.bodies(), "arbiters": gw.arbiters(),
               "cons": gw.constraints() if cons is not None else None}
        gw.step(dt)
        ref["keys"] = gw.solve_order()[0]
        ref["after"] = gw.bodies()
    start_bodies = gw.bodies()
    start_arbiters = gw.arbiters()
    start_cons = gw.constraints() if cons is not None else None
    con_order = gw.constraint_order()[0] if cons is not None else None
    n_cons = 0 if cons is None else int(np.sum(cons["enabled"] != 0))

    solve_ms, step_ms, bytes_per_launch, contacts = [], [], [], []
    stage_keys = ("ms_broadphase", "ms_narrowphase", "ms_epa", "ms_contact_cache", "ms_colouring", "ms_solve", "ms_update")
    stage_sum = dict.fromkeys(stage_keys, 0.0)
    first_step = gw.stats()["steps"]
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
    wall = t_end - t_start
    value = n_dyn * args.steps / wall
    ms_per_step = 1000.0 * wall / args.steps
    avg_solve_s = float(np.mean(solve_ms)) / 1000.0
    achieved = float(np.mean(bytes_per_launch)) / avg_solve_s / 1e9
    final = gw.stats()
    lws = None if args.no_lws else learned_warm_start_leg(gw, dt, cfg, max(args.lws_settle, gw.stats()["steps"]),
                                                          args.lws_iterations, 20)
    gw.close()

    cpu = None
    if not args.no_cpu:
        from tests import orc
        from tests.orc import OracleWorld

        threads = orc.set_threads(int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1))
        ow = OracleWorld(start_bodies, shapes, verts, cfg)
        ow.set_arbiters(start_arbiters)
        if start_cons is not None:
            ow.set_constraints(start_cons)
            ow.set_constraint_order(con_order)
        secs = ow.time_steps(dt, args.cpu_steps)
        del ow
        # FP32 error of one GPU step against the FP64 oracle (Jitter2.Double) replaying it from the same
        # state in the GPU's solve order
        o64 = OracleWorld(ref["bodies"], shapes, verts, cfg, fp64=True)
        o64.set_arbiters(ref["arbiters"])
        if ref["cons"] is not None:
            o64.set_constraints(ref["cons"])
            o64.set_constraint_order(con_order)
        o64.step(dt, ref["keys"])
        s64 = o64.state64()
        del o64
        g = ref["after"]
        dyn = bodies["motion_type"] == 0
        dp = np.abs(g["position"].astype(np.float64) - s64[:, 0:3])[dyn]
        dv = np.abs(g["velocity"].astype(np.float64) - s64[:, 7:10])[dyn]
        dw = np.abs(g["angular_velocity"].astype(np.float64) - s64[:, 10:13])[dyn]
        speed = np.linalg.norm(s64[dyn, 7:10], axis=1)
        fp64 = {"steps": 1, "bodies": int(dyn.sum()), "max_abs_dp_m": float(dp.max(initial=0.0)),
                "max_abs_dv_mps": float(dv.max(initial=0.0)), "max_abs_dw_radps": float(dw.max(initial=0.0)),
                "rms_dv_mps": float(np.sqrt(np.mean(dv ** 2))) if dv.size else 0.0,
                "max_speed_mps": float(speed.max(initial=0.0)),
                "what": "FP32 GPU step vs the FP64 oracle stepping the same state in the GPU's colour order"}
        orc.set_threads(1)
        cpu = {"value": n_dyn * args.cpu_steps / secs, "unit": "body-steps/s", "cores": threads, "kind": "port",
               "label": "restatement CPU (the reference, C#, cannot run here)",
               "sample": f"{args.cpu_steps} steps of the same {n_dyn}-body scene continued from the GPU state at "
                         f"the start of the timed region: the oracle (C++ restatement, FP32) on {threads} host "
                         f"threads, narrowphase, boxes, integration and body update in parallel, Gauss-Seidel "
                         f"solve sequential in key order ({secs:.1f} s)",
               "fp64_error": fp64}

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
                   "timed_steps": [first_step, first_step + args.steps],
                   "settle": args.settle,
                   "contacts_timed_mean": float(np.mean(contacts)),
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
