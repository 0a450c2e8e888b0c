"""Planner benchmark: batched persistent launch vs per-query launches on 1024 uneven 7-DOF queries.

A step is one batched plan of the whole 1024-query workload (one launch of the persistent
planner).  ``value`` is plans/s of the batched planner with every input already in HBM; the
same queries are also run as 1024 sequential single-query launches (each followed by a stream
synchronisation, as a per-query caller would) and both solve rates are reported.  FK+CC
throughput is measured on 10^6 configurations against the large-scene world.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--queries 1024] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector (datasheet)
PEAK_HBM_GBS = 8000.0
# HBM bytes of one batched step of the default 1024-query workload (plan_pass1 + plan_pass2),
# from rocprofv3 FETCH_SIZE (x 2.0, calibrated on a plain 1 GiB copy) and WRITE_SIZE (x 1.0).
TRAFFIC_BYTES_PER_STEP = 16.82e9

METRIC = "plans/sec + p50 ms/plan, batched vs per-query at equal solve rate; FK+CC cfg/s"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_c3(n_queries: int, checker_factory):
    """Robot, per-query worlds/goals/starts/max_nodes of the uneven 1024-query workload."""
    from prrax_amd import scenes

    robot = scenes.make_arm(0, 64, 200)
    worlds, starts, goals, nodes = [], [], [], []
    for qi in range(n_queries):
        w, ng, N = scenes.c3_query_world(robot, qi)
        cc = dict(robot.ctx)
        cc.update(w)
        free = checker_factory(cc)
        rng = np.random.default_rng(scenes.query_seed(3, qi))
        sg = scenes.sample_free(rng, robot, 1 + ng, free)
        worlds.append(cc)
        starts.append(sg[0])
        goals.append(sg[1:])
        nodes.append(N)
    return robot, worlds, starts, goals, nodes


def _cpu_plan(args):
    cc, start, goals, kw = args
    from oracle import oracle as O

    r = O.plan(O.Context(cc, 7), start, goals, **kw)
    return r["status"] == O.STATUS_SOLVED


def cpu_baseline(worlds, starts, goals, nodes, kw, n_sample: int, workers: int):
    jobs = [(worlds[i], starts[i], goals[i], dict(kw, max_nodes=nodes[i])) for i in range(n_sample)]
    t0 = time.perf_counter()
    # ctypes releases the GIL for the duration of each oracle call, so threads run in parallel.
    with ThreadPoolExecutor(max_workers=workers) as ex:
        solved = list(ex.map(_cpu_plan, jobs))
    dt = time.perf_counter() - t0
    return {"value": round(n_sample / dt, 3), "unit": "plans/s", "cores": workers, "kind": "port",
            "sample": f"first {n_sample} of the {len(worlds)} queries, C oracle planner (same lockstep algorithm), "
                      f"one query per thread",
            "solve_rate": round(float(np.mean(solved)), 4), "wall_s": round(dt, 2)}


def copy_calibration(mib: int):
    """One plain copy of ``mib`` MiB: its FETCH_SIZE / WRITE_SIZE readings calibrate the HBM counters."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copy calibration: {mib} MiB read + {mib} MiB written")


def scale_leg(n_queries: int, checker_factory, kw: dict, launch: dict):
    """The same uneven-query generator at a larger batch, one launch: the few queries that never
    connect bound a 1024-query step, and a larger batch spreads them over more plans."""
    import torch

    from prrax_amd import prepare_batch

    t0 = time.perf_counter()
    robot, worlds, starts, goals, nodes = build_c3(n_queries, checker_factory)
    log(f"scaling: {n_queries} queries built in {time.perf_counter() - t0:.1f}s")
    pb = prepare_batch(np.stack(starts), goals, collision_context=worlds, max_nodes=max(nodes),
                       max_nodes_per_query=nodes, **launch, **kw)
    pb.enqueue()
    torch.cuda.synchronize()
    t = time.perf_counter()
    pb.enqueue()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    h = pb.bufs.host_summary()
    out = {"queries": n_queries, "ms": round(dt * 1000, 1), "plans_per_s": round(n_queries / dt, 1),
           "solve_rate": round(float(np.mean(h["solved"] == 1)), 4),
           "longest_query_ms": round(float(h["kernel_time_ms"].max()), 1)}
    log(f"scaling: {out}")
    del pb
    return out


def dual_arm_leg(n_queries: int):
    """14-DOF dual arm with mimic grippers (128 spheres, ~2000 self pairs), 150 spheres + 12
    boxes + floor, goals 1-4 per query, step 0.3: one batched launch vs per-query launches."""
    import torch

    from prrax_amd import prepare_batch, scenes
    from prrax_amd.prrtc import CollisionChecker

    sc = scenes.scene_c4()
    cc, robot = sc.context(), sc.robot
    free = CollisionChecker(cc, robot.dim, collision_checker="differentiable", collision_margin=1e-5).is_free
    starts, goals = [], []
    for qi in range(n_queries):
        rng = np.random.default_rng(scenes.query_seed(4, qi))
        ng = int(rng.integers(1, 5))
        sg = scenes.sample_free(rng, robot, 1 + ng, free)
        starts.append(sg[0])
        goals.append(sg[1:])
    kw = dict(step_size=0.3, num_new_samples=128, max_iterations=10_000, max_nodes=32_768, min_vals=robot.lower,
              max_vals=robot.upper)
    pb = prepare_batch(np.stack(starts), goals, collision_context=cc, **kw)
    pb.enqueue()
    torch.cuda.synchronize()
    t = time.perf_counter()
    pb.enqueue()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    solved = pb.bufs.host_summary()["solved"] == 1
    singles = [prepare_batch(starts[i][None], [goals[i]], collision_context=cc, **kw) for i in range(n_queries)]
    singles[0].enqueue()
    torch.cuda.synchronize()
    times = []
    for sp in singles:
        t = time.perf_counter()
        sp.enqueue()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    solved_pq = np.array([sp.bufs.host_summary()["solved"][0] == 1 for sp in singles])
    out = {"workload": f"{n_queries} queries, 14-DOF dual arm (128 spheres, mimic fingers), 150 spheres + 12 boxes "
                       "+ floor, 1-4 goals, step 0.3, S 128, max_nodes 32768",
           "batched_plans_per_s": round(n_queries / dt, 2), "batched_ms": round(dt * 1000, 2),
           "solve_rate": round(float(solved.mean()), 4),
           "per_query_plans_per_s": round(n_queries / sum(times), 2),
           "per_query_p50_ms": round(1000 * float(np.median(times)), 3),
           "identical_solved_flags": bool(np.array_equal(solved, solved_pq))}
    log(f"dual arm: {out}")
    return out


def refine_leg(n_paths: int, waypoints: int = 64, steps: int = 50):
    """Path refinement: n_paths planner paths of the 7-DOF arm in a 100-sphere world, resampled
    to 64 waypoints, 50 gradient steps of the sphere-penetration cost (hand-written FK adjoint),
    each moved waypoint kept only if the exact predicate passes."""
    import torch

    from prrax_amd import prrtc_plan_batch, scenes
    from prrax_amd.prrtc import CollisionChecker
    from prrax_amd.refine import SphereCost, refine_paths, resample_path

    sc = scenes.scene_c5()
    cc, robot = sc.context(), sc.robot
    free = CollisionChecker(cc, robot.dim, collision_checker="differentiable", collision_margin=1e-5).is_free
    starts, goals = [], []
    for qi in range(n_paths):
        sg = scenes.sample_free(np.random.default_rng(scenes.query_seed(5, qi)), robot, 2, free)
        starts.append(sg[0])
        goals.append(sg[1:])
    res = prrtc_plan_batch(np.stack(starts), goals, collision_context=cc, max_iterations=10_000, max_nodes=16_384,
                           min_vals=robot.lower, max_vals=robot.upper)
    paths = [resample_path(r.path, waypoints) for r in res if r.solved]
    P = np.stack(paths)
    log(f"refine: {P.shape[0]}/{n_paths} planned paths, {waypoints} waypoints each")
    flops = torch.zeros(1, dtype=torch.int64, device="cuda")
    refine_paths(P, cc, iterations=1, step=2e-3, min_vals=robot.lower, max_vals=robot.upper)  # warm-up
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    out, costs = refine_paths(P, cc, iterations=steps, step=2e-3, min_vals=robot.lower, max_vals=robot.upper,
                              flops=flops)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    moved = float(np.mean(np.any(out.cpu().numpy() != P, axis=2)[:, 1:-1]))
    # cost + adjoint kernel alone, all waypoints of all paths in one launch
    skc = SphereCost(cc, robot.dim)
    q = torch.from_numpy(P.reshape(-1, robot.dim)).cuda()
    kfl = torch.zeros(1, dtype=torch.int64, device="cuda")
    skc(q, grad=True)
    torch.cuda.synchronize()
    reps = 10
    e0.record()
    for _ in range(reps):
        skc(q, grad=True, flops=kfl)
    e1.record()
    torch.cuda.synchronize()
    kms = e0.elapsed_time(e1) / reps
    kt = float(kfl.item()) / reps / (kms / 1000.0) / 1e12
    out = {"workload": f"{P.shape[0]} planner paths x {waypoints} waypoints, 7-DOF arm (48 spheres), 100 world "
                       f"spheres, {steps} steps, beta 50, smoothness 0.1",
           "paths": int(P.shape[0]), "ms": round(ms, 3), "paths_per_s": round(P.shape[0] / (ms / 1000.0), 1),
           "cost_first": round(float(costs[0]), 4), "cost_last": round(float(costs[-1]), 4),
           "interior_waypoints_moved": round(moved, 4),
           "cost_grad_kernel": {"configs": int(q.shape[0]), "ms": round(kms, 4),
                                "cfg_per_s": round(q.shape[0] / (kms / 1000.0), 1), "tflops": round(kt, 3),
                                "frac_of_fp32_peak": round(kt / PEAK_FP32_TFLOPS, 4)}}
    log(f"refine: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--queries", type=int, default=1024)
    ap.add_argument("--samples", type=int, default=128, help="explorers (samples per round) per query")
    ap.add_argument("--granularity", type=int, default=16)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--slots", type=int, default=0)
    ap.add_argument("--per-query", type=int, default=-1, help="queries run as single launches (-1 = all)")
    ap.add_argument("--fkcc-configs", type=int, default=1_000_000)
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--max-iterations", type=int, default=10_000)
    ap.add_argument("--team-max", type=int, default=0, help="largest work-group team per open query (0 = library default)")
    ap.add_argument("--refine-paths", type=int, default=512, help="paths of the refinement leg (0 = skip)")
    ap.add_argument("--dual-arm-queries", type=int, default=256, help="queries of the 14-DOF leg (0 = skip)")
    ap.add_argument("--scale-queries", type=int, default=4096,
                    help="queries of the batch-size scaling leg (same generator, one launch; 0 = skip)")
    ap.add_argument("--no-autotune", action="store_true", help="use S = --samples and the library's launch defaults")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every query is independent and runs on a single device; "
              "runs on several GPUs are not supported (the query queue would be split per GPU).")
        sys.exit(2)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return

    import torch

    from prrax_amd import prepare_batch, scenes
    from prrax_amd.prrtc import CollisionChecker

    torch.cuda.set_device(0)
    dev_name = torch.cuda.get_device_name(0)
    log(f"device {dev_name}")

    def checker_factory(cc):
        chk = CollisionChecker(cc, 7, collision_checker="differentiable", collision_margin=1e-5)
        return chk.is_free

    t0 = time.perf_counter()
    robot, worlds, starts, goals, nodes = build_c3(args.queries, checker_factory)
    log(f"workload built in {time.perf_counter() - t0:.1f}s")
    kw = dict(max_iterations=args.max_iterations, step_size=0.5, num_new_samples=args.samples, granularity=args.granularity,
              min_vals=robot.lower, max_vals=robot.upper)
    launch = dict(threads=args.threads, slots=args.slots, team_max=args.team_max)
    tuned = None
    if not args.no_autotune:
        from prrax_amd.autotune import autotune

        t0 = time.perf_counter()
        base = {k: v for k, v in kw.items() if k != "num_new_samples"}
        tuned = autotune(starts, goals, worlds, max_nodes_per_query=nodes, log=log, **base)
        kw["num_new_samples"] = tuned.num_new_samples
        launch.update({k: v for k, v in tuned.launch.items()})
        log(f"autotune: chose S={tuned.num_new_samples} {tuned.launch} in {time.perf_counter() - t0:.1f}s")
    # Queue order: the library's default, longest predicted first (results do not depend on it).
    pb = prepare_batch(np.stack(starts), goals, collision_context=worlds, max_nodes=max(nodes),
                       max_nodes_per_query=nodes, **launch, **kw)
    # ---- batched
    for w in range(args.warmup):
        t = time.perf_counter()
        pb.enqueue()
        torch.cuda.synchronize()
        log(f"warmup step {w}: {time.perf_counter() - t:.3f}s")
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        pb.enqueue()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps
    ms_per_step = wall * 1000.0 / args.steps
    summ = pb.bufs.host_summary()
    stats = pb.bufs.stats()
    solved_b = summ["solved"] == 1
    plans_per_s = args.queries / (ms_per_step / 1000.0)
    log(f"batched: {ms_per_step:.2f} ms/step ({kernel_ms:.2f} ms device), {plans_per_s:.1f} plans/s, "
        f"solve rate {solved_b.mean():.4f}, {stats}")
    # ---- per-query launches
    n_pq = args.queries if args.per_query < 0 else min(args.per_query, args.queries)
    pq = None
    if n_pq > 0:
        singles = [prepare_batch(starts[i][None], [goals[i]], collision_context=[worlds[i]], max_nodes=nodes[i],
                                 threads=launch["threads"], **kw) for i in range(n_pq)]
        singles[0].enqueue()
        torch.cuda.synchronize()
        times = []
        for k, sp in enumerate(singles):
            t = time.perf_counter()
            sp.enqueue()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
            if (k + 1) % 128 == 0:
                log(f"per-query launches: {k + 1}/{n_pq} done, {sum(times):.2f}s")
        solved_pq = np.array([sp.bufs.host_summary()["solved"][0] == 1 for sp in singles])
        pq_plans_per_s = n_pq / sum(times)
        same = bool(np.array_equal(solved_pq, solved_b[:n_pq]))
        log(f"per-query: {pq_plans_per_s:.1f} plans/s, p50 {1000 * np.median(times):.3f} ms, solve rate "
            f"{solved_pq.mean():.4f}, identical solved flags: {same}")
        pq = {"plans_per_s": round(pq_plans_per_s, 2), "p50_ms_per_plan": round(1000 * float(np.median(times)), 3),
              "solve_rate": round(float(solved_pq.mean()), 4), "queries": n_pq, "identical_solved_flags": same}
    # ---- FK + CC throughput (large-scene world, 10^6 configs)
    sc2 = scenes.scene_c2()
    rng = np.random.default_rng(11)
    qcfg = torch.from_numpy(rng.uniform(sc2.robot.lower, sc2.robot.upper,
                                        size=(args.fkcc_configs, 7)).astype(np.float32)).cuda()
    fkcc = {}
    # verdicts (binary; margin 0.005) use the coarse-culled check; the min-distance leg is the
    # flat full sweep over every sphere pair (no early exit, no culling)
    legs = (("binary", "binary", True), ("margin_verdict", "differentiable", True),
            ("min_distance_full_sweep", "differentiable", False))
    for label, mode, verdict in (legs if args.fkcc_configs > 0 else ()):
        # margin 0.005: the floor lies 0.01 m under the base spheres, so 0.02 flags every config
        chk = CollisionChecker(sc2.context(), 7, collision_checker=mode, collision_margin=0.005)
        hit = torch.empty(args.fkcc_configs, dtype=torch.uint8, device="cuda") if verdict else None
        mind = None if verdict else torch.empty(args.fkcc_configs, dtype=torch.float32, device="cuda")
        flops = torch.zeros(1, dtype=torch.int64, device="cuda")
        chk.enqueue(qcfg, hit, mind)
        torch.cuda.synchronize()
        flops.zero_()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 3
        e0.record()
        for _ in range(reps):
            chk.enqueue(qcfg, hit, mind, flops)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        fl = float(flops.item()) / reps
        fkcc[label] = {"cfg_per_s": round(args.fkcc_configs / (ms / 1000.0), 1), "ms": round(ms, 3),
                       "tflops": round(fl / (ms / 1000.0) / 1e12, 3),
                       "frac_of_fp32_peak": round(fl / (ms / 1000.0) / 1e12 / PEAK_FP32_TFLOPS, 4)}
        if verdict:
            fkcc[label]["in_collision"] = round(float(hit.float().mean().item()), 4)
        log(f"fkcc {label}: {fkcc[label]}")
    scale = scale_leg(args.scale_queries, checker_factory, kw, launch) if args.scale_queries > 0 else None
    refine = refine_leg(args.refine_paths) if args.refine_paths > 0 else None
    dual = dual_arm_leg(args.dual_arm_queries) if args.dual_arm_queries > 0 else None
    achieved = stats["flops"] / (kernel_ms / 1000.0) / 1e12
    default_workload = (args.queries == 1024 and kw["num_new_samples"] == 128 and args.granularity == 16
                        and args.max_iterations == 10_000)
    out = {
        "metric": METRIC,
        "value": round(plans_per_s, 2),
        "unit": "plans/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{args.queries} uneven 7-DOF queries (10-300 spheres, 0-24 boxes, floor, 1-8 goals, "
                               "max_nodes 8192-65536), one persistent launch",
                   "queries": args.queries, "num_new_samples": kw["num_new_samples"], "granularity": args.granularity,
                   "step_size": 0.5, "max_iterations": args.max_iterations, "launch": stats.get("launch"),
                   "autotuned": tuned is not None},
        "batched": {"plans_per_s": round(plans_per_s, 2), "solve_rate": round(float(solved_b.mean()), 4),
                    "p50_ms_per_plan_device": round(float(np.median(summ["kernel_time_ms"])), 3),
                    "kernel_ms_per_step": round(kernel_ms, 3), "configs_checked": stats["configs"],
                    "rounds": stats["rounds"]},
        "per_query": pq,
        "batched_over_per_query": round(plans_per_s / pq["plans_per_s"], 2) if pq else None,
        "fkcc": {"workload": f"{args.fkcc_configs} uniform configs vs large-scene world (64 spheres, 200 self pairs, "
                             "200 spheres, 16 boxes, 8 capsules, floor)", **fkcc},
        "autotune": None if tuned is None else {
            "chosen": tuned.kwargs(), "sample_plans_per_s": round(tuned.plans_per_s, 1),
            "candidates": [{k: (round(v, 2) if isinstance(v, float) else v) for k, v in r.items()} for r in tuned.table]},
        "batch_scaling": scale,
        "refine": refine,
        "dual_arm": dual,
        "roofline": {"kernel": "plan_pass1+plan_pass2", "bound": "valu", "achieved": round(achieved, 4),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 5),
                     "traffic": TRAFFIC_BYTES_PER_STEP if default_workload else None,
                     "flops_per_launch": stats["flops"]},
        "device": dev_name,
    }
    if not args.no_cpu:
        workers = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {args.cpu_sample} queries on {workers} threads")
        out["cpu_baseline"] = cpu_baseline(worlds, starts, goals, nodes, kw, min(args.cpu_sample, args.queries),
                                           workers)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
