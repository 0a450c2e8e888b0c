"""Planner benchmark: batched persistent launch vs per-query launches on 1024 uneven 7-DOF queries.

A step is one batched plan of the whole 1024-query workload (one launch of the persistent
planner: reset, per-query pass, cooperative team pass).  ``value`` is plans/s of the batched
planner with every input already in HBM.  The autotuner picks samples per round, waves per query
and the FK/check tile on a tail-seeded sample first.  Outside the timed region every returned
path is re-validated by the host FP64 checker at t = k/16, the same queries run as 1024
sequential single-query launches (each followed by a synchronisation, as a per-query caller
would), and the other legs run: FK+CC throughput on 10^6 configurations against the large-scene
world, the large scene itself (100 seeded queries, S = 512, 65536 nodes), a 4096-query batch,
the 14-DOF dual arm, path refinement, and the CPU baselines (the reference's host RRT-Connect
loop restated in NumPy, on the host's cores).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--queries 1024] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import multiprocessing
import os
import platform
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector (datasheet)
PEAK_HBM_GBS = 8000.0
# HBM bytes of one batched step of the 1024-query workload (plan_pass1 + plan_pass2): rocprofv3
# FETCH_SIZE (x 2.0, calibrated on a plain 1 GiB copy) plus WRITE_SIZE (x 1.0), measured by the
# profile runs named in "source", not by this run.  Reported only for the launch shape they ran.
TRAFFIC = {"bytes_per_step": 25.28e9,
           "source": "profiles/plan_batched_1024q_hbm_counters.md: rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 "
                     "of the timed step of the same workload at this launch shape (244.2 ms step)",
           "shape": {"num_new_samples": 512, "slots": 256, "threads": 512, "pass1_rounds": 24, "team_max": 512,
                     "fk_batch": 16}}

METRIC = "plans/sec + p50 ms/plan, batched vs per-query at equal solve rate; FK+CC cfg/s"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_workers() -> int:
    """Host threads this run may use: the process's CPU affinity, at most 16 (one GPU's share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def host_info() -> dict:
    model = platform.processor() or ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except (OSError, subprocess.SubprocessError, ValueError):
        nproc = os.cpu_count() or 0
    return {"nproc": nproc, "cpu_model": model, "affinity": cpu_workers()}


# ------------------------------------------------------------------------------ workloads
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


def gpu_free_checker(dim: int):
    from prrax_amd.prrtc import CollisionChecker

    def factory(cc):
        return CollisionChecker(cc, dim, collision_checker="differentiable", collision_margin=1e-5).is_free

    return factory


def time_singles(singles, label: str):
    import torch

    singles[0].enqueue()
    torch.cuda.synchronize()
    times = []
    for k, sp in enumerate(singles):
        t = time.perf_counter()
        sp.enqueue()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        if (k + 1) % 256 == 0:
            log(f"{label}: {k + 1}/{len(singles)} single-query launches, {sum(times):.2f}s")
    solved = np.array([sp.bufs.host_summary()["solved"][0] == 1 for sp in singles])
    return np.array(times), solved


def tail_leg(q: int, start, goals, world, max_nodes: int, kw: dict) -> dict:
    """The longest query of the per-query leg alone, on a team of every work-group (team cap 256)
    and on a team of 64 (its share while four long queries run at once in the batch): time, rounds
    and the device-clock time per round of each phase (the planner's phase counters, 100 MHz)."""
    import torch

    from prrax_amd import prepare_batch

    out = {"query": int(q), "world": {"spheres": int(len(world["world_spheres"])), "boxes": int(len(world["world_boxes"]))},
           "max_nodes": int(max_nodes), "teams": {}}
    for tm in (256, 64):
        sp = prepare_batch(start[None], [goals], collision_context=[world], max_nodes=max_nodes, team_max=tm, **kw)
        sp.enqueue()
        torch.cuda.synchronize()
        t = time.perf_counter()
        sp.enqueue()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        st, h = sp.bufs.stats(), sp.bufs.host_summary()
        r = max(int(h["iterations"][0]), 1)
        out["teams"][str(tm)] = {"ms": round(1000 * dt, 2), "rounds": r, "us_per_round": round(1e6 * dt / r, 1),
                                 "status": int(h["status"][0]), "tree_sizes": [int(v) for v in h["sizes"][0]],
                                 "phase_us_per_round": {k: round(v / r / 100.0, 2) for k, v in st["pass2_ticks"].items()}}
        del sp
    log(f"tail query: {out}")
    return out


def validate_leg(results, worlds) -> dict:
    """Every returned path re-validated by the host FP64 checker at t = k/16 of each edge."""
    from prrax_amd.validate import validate_paths

    t0 = time.perf_counter()
    paths = [r.path if r.solved else None for r in results]
    ok = validate_paths(worlds, paths, resolution=16, tol=1e-5, workers=cpu_workers())
    solved = int(sum(r.solved for r in results))
    out = {"solved": solved, "validated": int(ok.sum()), "checker": "host FP64 signed distances, t = k/16 of every "
           "edge, contact tolerance 1e-5 m", "wall_s": round(time.perf_counter() - t0, 2)}
    log(f"validation: {out}")
    return out


def c2_leg(n_queries: int, kw_base: dict):
    """The large scene (64 spheres, 200 self pairs, 200 spheres, 16 boxes, 8 capsules, floor; 4
    goals, S = 512, 65536 nodes): one batched launch of the seeded queries and per-query launches."""
    import torch

    from prrax_amd import prepare_batch, scenes

    sc = scenes.scene_c2()
    cc = sc.context()
    free = gpu_free_checker(7)(cc)
    starts, goals = [], []
    for qi in range(n_queries):
        sg = scenes.sample_free(np.random.default_rng(scenes.query_seed(2, qi)), sc.robot, 5, free)
        starts.append(sg[0])
        goals.append(sg[1:])
    kw = dict(kw_base, num_new_samples=512, max_nodes=65536, min_vals=sc.robot.lower, max_vals=sc.robot.upper)
    pb = prepare_batch(np.stack(starts), np.stack(goals), collision_context=cc, **kw)
    pb.enqueue()
    torch.cuda.synchronize()
    t = time.perf_counter()
    pb.enqueue()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    h = pb.bufs.host_summary()
    singles = [prepare_batch(starts[i][None], goals[i][None], collision_context=cc, **kw) for i in range(n_queries)]
    times, solved_pq = time_singles(singles, "large scene")
    out = {"workload": f"{n_queries} seeded queries, large scene (64 spheres, 200 self pairs, 200 spheres, 16 boxes, "
                       "8 capsules, floor), 4 goals, S 512, max_nodes 65536",
           "batched_plans_per_s": round(n_queries / dt, 1), "batched_ms": round(1000 * dt, 2),
           "p50_ms_per_plan_device": round(float(np.median(h["kernel_time_ms"])), 3),
           "solve_rate": round(float(np.mean(h["solved"] == 1)), 4),
           "per_query_plans_per_s": round(n_queries / times.sum(), 1),
           "per_query_p50_ms": round(1000 * float(np.median(times)), 3),
           "identical_solved_flags": bool(np.array_equal(h["solved"] == 1, solved_pq))}
    log(f"large scene: {out}")
    return out


def scale_leg(n_queries: int, kw: dict, launch: dict):
    """The same uneven-query generator at a larger batch, one launch: the few queries that never
    connect bound a 1024-query step, and a larger batch spreads them over more plans."""
    import torch

    from prrax_amd import prepare_batch

    t0 = time.perf_counter()
    robot, worlds, starts, goals, nodes = build_c3(n_queries, gpu_free_checker(7))
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

    sc = scenes.scene_c4()
    cc, robot = sc.context(), sc.robot
    free = gpu_free_checker(robot.dim)(cc)
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
    times, solved_pq = time_singles(singles, "dual arm")
    out = {"workload": f"{n_queries} queries, 14-DOF dual arm (128 spheres, mimic fingers), 150 spheres + 12 boxes "
                       "+ floor, 1-4 goals, step 0.3, S 128, max_nodes 32768",
           "batched_plans_per_s": round(n_queries / dt, 2), "batched_ms": round(dt * 1000, 2),
           "solve_rate": round(float(solved.mean()), 4),
           "per_query_plans_per_s": round(n_queries / times.sum(), 2),
           "per_query_p50_ms": round(1000 * float(np.median(times)), 3),
           "identical_solved_flags": bool(np.array_equal(solved, solved_pq))}
    log(f"dual arm: {out}")
    return out


def refine_leg(n_paths: int, waypoints: int = 64, steps: int = 50):
    """Path refinement: n_paths planner paths of the 7-DOF arm in a 100-sphere world, resampled
    to 64 waypoints, 50 gradient steps of the sphere-penetration cost (hand-written FK adjoint),
    each move kept only if the exact predicate passes at the waypoint and along both edges."""
    import torch

    from prrax_amd import prrtc_plan_batch, scenes
    from prrax_amd.refine import SphereCost, refine_paths, resample_path

    sc = scenes.scene_c5()
    cc, robot = sc.context(), sc.robot
    free = gpu_free_checker(robot.dim)(cc)
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


def fkcc_leg(n_cfg: int) -> dict:
    """FK+CC throughput: n_cfg uniform configurations against the large-scene world.  Verdicts
    (binary; margin 0.005) use the coarse-culled check; the min-distance leg is the flat full
    sweep over every sphere pair (no early exit, no culling)."""
    import torch

    from prrax_amd import scenes
    from prrax_amd.prrtc import CollisionChecker

    sc2 = scenes.scene_c2()
    rng = np.random.default_rng(11)
    qcfg = torch.from_numpy(rng.uniform(sc2.robot.lower, sc2.robot.upper, size=(n_cfg, 7)).astype(np.float32)).cuda()
    fkcc = {}
    legs = (("binary", "binary", True), ("margin_verdict", "differentiable", True),
            ("min_distance_full_sweep", "differentiable", False))
    for label, mode, verdict in legs:
        # margin 0.005: the floor lies 0.01 m under the base spheres, so 0.02 flags every config
        chk = CollisionChecker(sc2.context(), 7, collision_checker=mode, collision_margin=0.005)
        hit = torch.empty(n_cfg, dtype=torch.uint8, device="cuda") if verdict else None
        mind = None if verdict else torch.empty(n_cfg, dtype=torch.float32, device="cuda")
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
        fkcc[label] = {"cfg_per_s": round(n_cfg / (ms / 1000.0), 1), "ms": round(ms, 3),
                       "tflops": round(fl / (ms / 1000.0) / 1e12, 3),
                       "frac_of_fp32_peak": round(fl / (ms / 1000.0) / 1e12 / PEAK_FP32_TFLOPS, 4)}
        if verdict:
            fkcc[label]["in_collision"] = round(float(hit.float().mean().item()), 4)
        log(f"fkcc {label}: {fkcc[label]}")
    return {"workload": f"{n_cfg} uniform configs vs large-scene world (64 spheres, 200 self pairs, 200 spheres, "
                        "16 boxes, 8 capsules, floor)", **fkcc}


def cloud_edges(sc, n_edges: int, rng) -> np.ndarray:
    """Edges [n_edges, 16, 7]: random starts in the limits, steps of about 0.3 rad."""
    a = rng.uniform(sc.robot.lower, sc.robot.upper, size=(n_edges, 1, 7))
    b = a + rng.normal(scale=0.3 / np.sqrt(7.0), size=a.shape)
    t = np.linspace(0.0, 1.0, 16)[None, :, None]
    return (a * (1.0 - t) + b * t).astype(np.float32)


def cloud_queries(n_queries: int, chk, sc, pcl):
    from prrax_amd import scenes

    def free(q):
        return chk.edges_free(np.asarray(q, np.float32)[:, None, :], sc.world, pcl, sc.extra["r_env"]).cpu().numpy() == 1

    return [scenes.sample_free(np.random.default_rng(scenes.query_seed(7, qi)), sc.robot, 2, free)
            for qi in range(n_queries)]


def cloud_leg(n_points: int, n_queries: int, n_edges: int = 65536) -> dict:
    """Point-cloud path: the grid build, edge validation against an n_points cloud (the points of
    the C1 world's spheres' surfaces) plus the floor, and the host-driven planner."""
    import torch

    from prrax_amd import PointCloud, PointCloudChecker, prrtc_plan_robogpu, scenes

    sc = scenes.scene_cloud(n_points)
    pts, r_env = sc.extra["points"], sc.extra["r_env"]
    chk = PointCloudChecker(sc.context(), 7)
    pcl = torch.from_numpy(pts).cuda()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    PointCloud(pcl, r_env, chk.max_radius)
    torch.cuda.synchronize()
    e0.record()
    grid = PointCloud(pcl, r_env, chk.max_radius)
    e1.record()
    torch.cuda.synchronize()
    build_ms = e0.elapsed_time(e1)
    rng = np.random.default_rng(13)
    wp = torch.from_numpy(cloud_edges(sc, n_edges, rng)).cuda()
    flops = torch.zeros(1, dtype=torch.int64, device="cuda")
    chk.edges_free(wp, sc.world, pcl, r_env)
    torch.cuda.synchronize()
    reps = 3
    e0.record()
    for _ in range(reps):
        out = chk.edges_free(wp, sc.world, pcl, r_env, flops)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    free_frac = float(out.float().mean().item())
    tf = float(flops.item()) / reps / (ms / 1000.0) / 1e12
    queries = cloud_queries(n_queries, chk, sc, pcl)
    kw = dict(max_iterations=2000, step_size=0.5, num_new_samples=128, granularity=16, max_nodes=100_000,
              connect_max_steps=32)
    times, solved, kernel_ms = [], [], []
    for qi, sg in enumerate(queries):
        t0 = time.perf_counter()
        r = prrtc_plan_robogpu(sg[0], sg[1:], sc.robot, chk, world_geom=sc.world, point_cloud=pcl, r_env=r_env,
                               seed=qi, **kw)
        times.append(time.perf_counter() - t0)
        solved.append(r.solved)
        kernel_ms.append(r.kernel_time_ms)
    out = {"workload": f"C1 arm (48 spheres) vs a {n_points}-point cloud (surfaces of 40 spheres, r_env {r_env}) "
                       f"+ floor; {n_edges} edges x 16 waypoints; {n_queries} host-driven plans "
                       f"(S {kw['num_new_samples']}, G 16, max_iterations {kw['max_iterations']})",
           "grid": {"cells": int(grid.struct.nx * grid.struct.ny * grid.struct.nz), "cell_m": round(grid.cell, 4),
                    "build_ms": round(build_ms, 3)},
           "check_edges": {"ms": round(ms, 3), "edges_per_s": round(n_edges / (ms / 1000.0), 1),
                           "cfg_per_s": round(16 * n_edges / (ms / 1000.0), 1), "tflops": round(tf, 3),
                           "free_fraction": round(free_frac, 4)},
           "plan": {"plans_per_s": round(len(times) / sum(times), 3), "p50_ms_per_plan": round(1000 * float(np.median(times)), 1),
                    "solve_rate": round(float(np.mean(solved)), 4),
                    "checker_share": round(float(sum(kernel_ms)) / (1000.0 * sum(times)), 4)}}
    log(f"cloud: {out}")
    return out, (sc, queries, kw)


def _cloud_cpu_job(job):
    from oracle import host_rrtc

    cc, pts, r_env, start, goals, kw, limit = job
    t = time.perf_counter()
    r = host_rrtc.plan(host_rrtc.CloudChecker(cc, len(start), pts, r_env), start, goals, time_limit=limit, **kw)
    return {"solved": bool(r["solved"]), "timed_out": bool(r["timed_out"]), "s": time.perf_counter() - t}


def cpu_cloud_leg(pool, workers: int, setup, n: int, limit: float) -> dict:
    """The same host loop with a CPU checker (float32 NumPy FK + a scipy k-d tree over the
    points), on the first n of the GPU leg's queries, one per worker process."""
    sc, queries, kw = setup
    cc = sc.context()
    jobs = [(cc, sc.extra["points"], sc.extra["r_env"], sg[0], sg[1:],
             dict(kw, min_vals=sc.robot.lower, max_vals=sc.robot.upper, seed=qi), limit)
            for qi, sg in enumerate(queries[:n])]
    t0 = time.perf_counter()
    res = pool.map(_cloud_cpu_job, jobs, chunksize=1)
    dt = time.perf_counter() - t0
    return {"value": round(len(jobs) / dt, 3), "unit": "plans/s", "cores": workers, "queries": len(jobs),
            "p50_s_per_plan": round(float(np.median([r["s"] for r in res])), 3),
            "solve_rate": round(float(np.mean([r["solved"] for r in res])), 4),
            "time_limited": int(sum(r["timed_out"] for r in res)),
            "checker": "float32 NumPy FK + scipy cKDTree nearest point per robot sphere", "wall_s": round(dt, 2)}


# ------------------------------------------------------------------------------ CPU baselines
def _one_thread():
    from threadpoolctl import threadpool_limits

    threadpool_limits(1)


def _host_plan_job(job):
    """One query of the reference's host RRT-Connect loop (NumPy restatement); runs in a worker
    process forked before the GPU was touched."""
    from oracle import host_rrtc

    cc, start, goals, kw, limit = job
    t = time.perf_counter()
    r = host_rrtc.plan(host_rrtc.BatchChecker(cc, len(start)), start, goals, time_limit=limit, **kw)
    return {"solved": bool(r["solved"]), "timed_out": bool(r["timed_out"]), "s": time.perf_counter() - t}


def cpu_host_loop_leg(pool, workers: int, worlds, starts, goals, nodes, kw, sample, limit: float) -> dict:
    jobs = [(worlds[i], starts[i], goals[i],
             dict(max_iterations=kw["max_iterations"], step_size=kw["step_size"], num_new_samples=128,
                  granularity=kw["granularity"], max_nodes=nodes[i], min_vals=kw["min_vals"], max_vals=kw["max_vals"],
                  seed=int(i)), limit) for i in sample]
    t0 = time.perf_counter()
    res = pool.map(_host_plan_job, jobs, chunksize=1)
    dt = time.perf_counter() - t0
    solved = [r["solved"] for r in res]
    return {"value": round(len(jobs) / dt, 3), "unit": "plans/s", "cores": workers, "kind": "port",
            "sample": f"{len(jobs)} of the {len(worlds)} queries (the queries still open after the probe's 24 rounds, "
                      f"then every {max(1, len(worlds) // max(1, len(jobs)))}th), the reference's host "
                      "RRT-Connect loop (prrtc_robogpu.py:183-264) with batched float32 NumPy FK + collision, "
                      f"S 128, G 16, one query per worker process, {limit:.0f} s wall limit per query",
            "solve_rate": round(float(np.mean(solved)), 4),
            "time_limited": int(sum(r["timed_out"] for r in res)),
            "p50_s_per_plan": round(float(np.median([r["s"] for r in res])), 3), "wall_s": round(dt, 2)}


def cpu_single_query_leg(pool, n: int) -> dict:
    """The plain single query on the CPU path (7-DOF arm, 48 spheres, 60 self pairs, 40 world
    spheres + floor, S 128, 16384 nodes): n seeded queries, one process each, p50 ms per plan."""
    from prrax_amd import scenes
    from prrax_amd.validate import min_distance

    sc = scenes.scene_c1()
    cc = sc.context()
    jobs = []
    for qi in range(n):
        rng = np.random.default_rng(scenes.query_seed(1, qi))
        sg = scenes.sample_free(rng, sc.robot, 2, lambda q: min_distance(cc, q) > 1e-5)
        jobs.append((cc, sg[0], sg[1:], dict(max_iterations=10_000, step_size=0.5, num_new_samples=128, granularity=16,
                                             max_nodes=16384, min_vals=sc.robot.lower, max_vals=sc.robot.upper,
                                             seed=qi), 60.0))
    res = pool.map(_host_plan_job, jobs, chunksize=1)
    s = [r["s"] for r in res]
    return {"workload": f"{n} single queries, 7-DOF arm (48 spheres, 60 self pairs), 40 world spheres + floor, S 128, "
                        "max_nodes 16384, one thread each (host RRT-Connect loop, NumPy)",
            "p50_ms_per_plan": round(1000 * float(np.median(s)), 1), "solve_rate": round(float(np.mean(
                [r["solved"] for r in res])), 4)}


def _lockstep_job(args):
    cc, start, goals, kw = args
    from oracle import oracle as O

    r = O.plan(O.Context(cc, 7), start, goals, **kw)
    return r["status"] == O.STATUS_SOLVED


def cpu_lockstep_leg(worlds, starts, goals, nodes, kw, sample, workers: int) -> dict:
    """The C restatement of the kernel's lockstep algorithm, one query per thread (ctypes releases
    the GIL for each oracle call)."""
    jobs = [(worlds[i], starts[i], goals[i], dict(kw, max_nodes=nodes[i])) for i in sample]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        solved = list(ex.map(_lockstep_job, jobs))
    dt = time.perf_counter() - t0
    return {"value": round(len(jobs) / dt, 3), "unit": "plans/s", "cores": workers,
            "sample": f"{len(jobs)} evenly spaced queries (not the probe's tail), the C restatement of the "
                      "kernels' lockstep algorithm",
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


# ------------------------------------------------------------------------------ main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--queries", type=int, default=1024)
    ap.add_argument("--samples", type=int, default=128, help="explorers per round when not autotuned")
    ap.add_argument("--granularity", type=int, default=16)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--slots", type=int, default=0)
    ap.add_argument("--team-max", type=int, default=0)
    ap.add_argument("--fk-batch", type=int, default=0)
    ap.add_argument("--per-query", type=int, default=-1, help="queries run as single launches (-1 = all)")
    ap.add_argument("--fkcc-configs", type=int, default=1_000_000)
    ap.add_argument("--max-iterations", type=int, default=10_000)
    ap.add_argument("--c2-queries", type=int, default=100, help="queries of the large-scene leg (0 = skip)")
    ap.add_argument("--refine-paths", type=int, default=512, help="paths of the refinement leg (0 = skip)")
    ap.add_argument("--dual-arm-queries", type=int, default=256, help="queries of the 14-DOF leg (0 = skip)")
    ap.add_argument("--scale-queries", type=int, default=4096,
                    help="queries of the batch-size scaling leg (same generator, one launch; 0 = skip)")
    ap.add_argument("--cloud-points", type=int, default=100_000, help="points of the point-cloud leg (0 = skip)")
    ap.add_argument("--cloud-queries", type=int, default=16, help="host-driven plans of the point-cloud leg")
    ap.add_argument("--no-validate", action="store_true", help="skip the FP64 re-validation of the paths")
    ap.add_argument("--no-autotune", action="store_true", help="use S = --samples and the library's launch defaults")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=32, help="queries of the host-loop CPU baseline")
    ap.add_argument("--cpu-limit", type=float, default=20.0, help="wall-time limit per CPU-baseline query (s)")
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

    # The CPU-baseline workers are forked now, before anything touches the GPU; each runs one query
    # at a time on one thread (their BLAS pools are limited to a thread each).
    workers = cpu_workers()
    pool = None if args.no_cpu else multiprocessing.get_context("fork").Pool(workers, initializer=_one_thread)

    import torch

    from prrax_amd import prepare_batch

    torch.cuda.set_device(0)
    dev_name = torch.cuda.get_device_name(0)
    log(f"device {dev_name}")
    t0 = time.perf_counter()
    robot, worlds, starts, goals, nodes = build_c3(args.queries, gpu_free_checker(7))
    log(f"workload built in {time.perf_counter() - t0:.1f}s")
    kw = dict(max_iterations=args.max_iterations, step_size=0.5, num_new_samples=args.samples,
              granularity=args.granularity, min_vals=robot.lower, max_vals=robot.upper)
    launch = dict(threads=args.threads, slots=args.slots, team_max=args.team_max, fk_batch=args.fk_batch)
    tuned = None
    if not args.no_autotune:
        from prrax_amd.autotune import autotune

        t0 = time.perf_counter()
        base = {k: v for k, v in kw.items() if k != "num_new_samples"}
        tuned = autotune(starts, goals, worlds, max_nodes_per_query=nodes, log=log, **base)
        kw["num_new_samples"] = tuned.num_new_samples
        launch.update(tuned.launch)
        log(f"autotune: chose S={tuned.num_new_samples} {tuned.launch} in {time.perf_counter() - t0:.1f}s")
    pb = prepare_batch(np.stack(starts), goals, collision_context=worlds, max_nodes=max(nodes),
                       max_nodes_per_query=nodes, **launch, **kw)
    # ---- batched
    for w in range(args.warmup):
        t = time.perf_counter()
        pb.enqueue()
        torch.cuda.synchronize()
        log(f"warmup step {w}: {time.perf_counter() - t:.3f}s")
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
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
    results = pb.bufs.results()
    solved_b = summ["solved"] == 1
    plans_per_s = args.queries / (ms_per_step / 1000.0)
    log(f"batched: {ms_per_step:.2f} ms/step ({kernel_ms:.2f} ms device), {plans_per_s:.1f} plans/s, "
        f"solve rate {solved_b.mean():.4f}, {stats}")
    validation = None if args.no_validate else validate_leg(results, worlds)
    # ---- per-query launches
    n_pq = args.queries if args.per_query < 0 else min(args.per_query, args.queries)
    pq = None
    if n_pq > 0:
        pq_launch = {k: v for k, v in launch.items() if k in ("threads", "fk_batch")}
        singles = [prepare_batch(starts[i][None], [goals[i]], collision_context=[worlds[i]], max_nodes=nodes[i],
                                 **pq_launch, **kw) for i in range(n_pq)]
        times, solved_pq = time_singles(singles, "per-query")
        same = bool(np.array_equal(solved_pq, solved_b[:n_pq]))
        pq = {"plans_per_s": round(n_pq / times.sum(), 2), "p50_ms_per_plan": round(1000 * float(np.median(times)), 3),
              "solve_rate": round(float(solved_pq.mean()), 4), "queries": n_pq, "identical_solved_flags": same,
              # the batched step cannot end before its longest query, which the per-query launch
              # already runs with the whole GPU: batched / per-query <= sum / max of these times
              "ratio_bound_sum_over_max": round(float(times.sum() / times.max()), 2),
              "longest_ms": round(1000 * float(times.max()), 1)}
        log(f"per-query: {pq}")
        del singles
        qt = int(np.argmax(times))
        pq["tail"] = tail_leg(qt, starts[qt], goals[qt], worlds[qt], nodes[qt], kw)
    fkcc = fkcc_leg(args.fkcc_configs) if args.fkcc_configs > 0 else None
    c2 = c2_leg(args.c2_queries, {k: v for k, v in kw.items() if k in ("max_iterations", "step_size", "granularity")}) \
        if args.c2_queries > 0 else None
    scale = scale_leg(args.scale_queries, kw, launch) if args.scale_queries > 0 else None
    refine = refine_leg(args.refine_paths) if args.refine_paths > 0 else None
    dual = dual_arm_leg(args.dual_arm_queries) if args.dual_arm_queries > 0 else None
    cloud, cloud_setup = cloud_leg(args.cloud_points, args.cloud_queries) if args.cloud_points > 0 else (None, None)
    # algorithmic FP32 work of the step: FK + collision tests performed, plus nearest-neighbour
    # scoring (3 * dim flops per (query, node) pair: a subtraction and an FMA per dimension)
    nn_flops = 3 * 7 * stats["nn_pairs"]
    achieved = (stats["flops"] + nn_flops) / (kernel_ms / 1000.0) / 1e12
    shape_now = dict(stats.get("launch", {}), num_new_samples=kw["num_new_samples"])
    same_shape = args.queries == 1024 and all(shape_now.get(k) == v for k, v in TRAFFIC["shape"].items())
    nn_gbs = stats["nn_bytes"] / (kernel_ms / 1000.0) / 1e9
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
                   "fk_batch": launch.get("fk_batch", 0), "autotuned": tuned is not None},
        "batched": {"plans_per_s": round(plans_per_s, 2), "solve_rate": round(float(solved_b.mean()), 4),
                    "p50_ms_per_plan_device": round(float(np.median(summ["kernel_time_ms"])), 3),
                    "kernel_ms_per_step": round(kernel_ms, 3), "configs_checked": stats["configs"],
                    "rounds": stats["rounds"], "nn_bytes_per_step": stats["nn_bytes"],
                    "nn_tile_passes": stats["nn_tile_passes"]},
        "validation": validation,
        "per_query": pq,
        "batched_over_per_query": round(plans_per_s / pq["plans_per_s"], 2) if pq else None,
        "fkcc": fkcc,
        "large_scene": c2,
        "autotune": None if tuned is None else {
            "chosen": tuned.kwargs(), "sample_plans_per_s": round(tuned.plans_per_s, 1),
            "sample_queries": len(tuned.sample), "tail_after_probe": len(tuned.tail),
            "tail_in_sample": len(set(tuned.sample) & set(tuned.tail)),
            "candidates": [{k: (round(v, 2) if isinstance(v, float) else v) for k, v in r.items()} for r in tuned.table]},
        "batch_scaling": scale,
        "refine": refine,
        "dual_arm": dual,
        "point_cloud": cloud,
        "roofline": {"kernel": "plan_pass1+plan_pass2", "bound": "valu", "achieved": round(achieved, 4),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 5),
                     "traffic": TRAFFIC["bytes_per_step"] if same_shape else None,
                     "traffic_source": TRAFFIC["source"] if same_shape else
                     f"not measured for this launch shape (profiled: {TRAFFIC['shape']})",
                     "flops_per_launch": stats["flops"] + nn_flops, "fkcc_flops_per_launch": stats["flops"],
                     "nn_flops_per_launch": nn_flops, "nn_bytes_per_launch": stats["nn_bytes"],
                     "nn_achieved_gbs": round(nn_gbs, 2), "nn_frac_of_hbm": round(nn_gbs / PEAK_HBM_GBS, 5)},
        "device": dev_name,
    }
    if pool is not None:
        info = host_info()
        order = list(tuned.tail) if tuned is not None else []
        stride = max(1, args.queries // max(1, args.cpu_sample - min(8, len(order))))
        sample = list(dict.fromkeys(order[:8] + list(range(0, args.queries, stride))))[:args.cpu_sample]
        log(f"cpu baseline: host loop on {len(sample)} queries, {workers} processes ({info})")
        cpu = cpu_host_loop_leg(pool, workers, worlds, starts, goals, nodes, kw, sample, args.cpu_limit)
        cpu.update(info)
        log(f"cpu baseline: {cpu}")
        out["cpu_baseline"] = cpu
        out["cpu_single_query"] = cpu_single_query_leg(pool, 8)
        log(f"cpu single query: {out['cpu_single_query']}")
        # the lockstep oracle runs the kernels' own algorithm; the tail queries (trees filled over
        # thousands of rounds) would take minutes each, so it times the evenly spaced part only
        even = [i for i in sample if i not in set(order)]
        out["cpu_lockstep_oracle"] = cpu_lockstep_leg(worlds, starts, goals, nodes, kw, even, workers)
        log(f"cpu lockstep oracle: {out['cpu_lockstep_oracle']}")
        if cloud_setup is not None:
            out["cpu_point_cloud"] = cpu_cloud_leg(pool, workers, cloud_setup, min(workers, args.cloud_queries),
                                                   args.cpu_limit)
            log(f"cpu point cloud: {out['cpu_point_cloud']}")
        pool.close()
        pool.join()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
