"""Benchmark-suite harness in the style of the reference's ``evaluate_mbm.py``.

The reference evaluates per-query ``prrtc_plan`` on the MotionBenchMaker problem sets
(``evaluate_mbm.py:285-437``): it verifies the robot's context once (``evaluate_mbm.py:70-224``),
then per valid problem builds the context with that problem's obstacles, plans, times the call
on the host (dispatch + kernel + synchronisation, ``evaluate_mbm.py:395-406``) and reports
solved / valid / total with time, iteration and path-cost statistics over the solved problems
(``evaluate_mbm.py:500-540``).  The MBM files (problems.pkl, URDFs) are not available offline;
this harness takes any robot description plus a problem list, in the same shape:

    problems = {"set name": [{"valid": True, "start": [...], "goals": [[...], ...],
                              "obstacles": [{"type": "sphere", "center": ..., "radius": ...}, ...]},
                             ...]}

Extensions: every returned path is also re-validated with the host FP64 checker at t = k/16
(``prrax_amd.validate``), and ``batched=True`` plans each problem set in one persistent launch.
"""

from __future__ import annotations

import json
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import robot_model as RM
from .validate import validate_paths

PERCENTILES = (25, 50, 75, 95)


@dataclass
class SuiteResult:
    rows: list = field(default_factory=list)  # one dict per solved problem
    total: int = 0
    valid: int = 0
    failed: int = 0
    path_invalid: int = 0

    @property
    def solved(self) -> int:
        return self.valid - self.failed


def describe(values) -> dict:
    """mean, std, min, 25/50/75/95 % and max of a column (the reference's ``describe``)."""
    v = np.asarray(values, np.float64)
    if v.size == 0:
        return {}
    out = {"mean": float(v.mean()), "std": float(v.std(ddof=1)) if v.size > 1 else 0.0, "min": float(v.min())}
    for p in PERCENTILES:
        out[f"{p}%"] = float(np.percentile(v, p))
    out["max"] = float(v.max())
    return out


def summary(res: SuiteResult) -> dict:
    rows = res.rows
    t = [r["planning_time_us"] for r in rows]
    return {"solved": res.solved, "valid": res.valid, "total": res.total, "paths_invalid": res.path_invalid,
            "planning_time_us": describe(t), "kernel_time_us": describe([r["kernel_time_us"] for r in rows]),
            "planning_iterations": describe([r["iterations"] for r in rows]),
            "path_cost": describe([r["cost"] for r in rows])}


def markdown_table(stats: dict, title: str) -> str:
    keys = ["mean", "std", "min", *[f"{p}%" for p in PERCENTILES], "max"]
    cols = [c for c in ("planning_time_us", "kernel_time_us", "planning_iterations", "path_cost") if stats.get(c)]
    lines = [f"| {title} | " + " | ".join(cols) + " |", "| --- |" + " --- |" * len(cols)]
    for k in keys:
        lines.append(f"| {k} | " + " | ".join(f"{stats[c][k]:.6g}" for c in cols) + " |")
    return "\n".join(lines)


def evaluate_suite(urdf: str, srdf: Optional[str], problems: dict, *, batched: bool = False, warmup: bool = True,
                   max_problems_per_set: int = 0, validate: bool = True, log=None, **plan_kw) -> SuiteResult:
    """Plan every valid problem of every set; ``plan_kw`` are ``prrtc_plan`` keywords
    (defaults: the limits of the URDF, collision_checker ``binary_coarse_coop`` as the
    reference's harness uses)."""
    import torch

    from .prrtc import prrtc_plan, prrtc_plan_batch

    base, lo, hi, model = RM.load_robot(urdf, srdf)
    RM.verify_collision_context(base, lo, hi, name=model.name)
    kw = dict(min_vals=lo, max_vals=hi, collision_checker="binary_coarse_coop")
    kw.update(plan_kw)
    res = SuiteResult()
    verified = False
    did_warmup = not warmup
    for set_name, pset in problems.items():
        items = pset if max_problems_per_set <= 0 else pset[:max_problems_per_set]
        valid_items = []
        for p in items:
            res.total += 1
            if not p.get("valid", False):
                continue
            res.valid += 1
            cc = dict(base, **RM.world_arrays(p.get("obstacles", ())))
            start = np.asarray(p["start"], np.float32)
            goals = np.asarray(p["goals"], np.float32).reshape(-1, start.shape[0])
            if not verified:
                RM.verify_collision_context(cc, lo, hi, start=start, goals=goals, name=model.name)
                verified = True
            valid_items.append((cc, start, goals))
        if not valid_items:
            continue
        if not did_warmup:
            cc, s, g = valid_items[0]
            prrtc_plan(s, g, collision_context=cc, **kw)
            torch.cuda.synchronize()
            did_warmup = True
        if batched:
            t0 = time.perf_counter_ns()
            results = prrtc_plan_batch(np.stack([s for _, s, _ in valid_items]), [g for _, _, g in valid_items],
                                       collision_context=[cc for cc, _, _ in valid_items], **kw)
            torch.cuda.synchronize()
            per = (time.perf_counter_ns() - t0) / len(valid_items)
            host_ns = [per] * len(valid_items)
        else:
            results, host_ns = [], []
            for cc, s, g in valid_items:
                t0 = time.perf_counter_ns()
                r = prrtc_plan(s, g, collision_context=cc, **kw)
                torch.cuda.synchronize()
                host_ns.append(time.perf_counter_ns() - t0)
                results.append(r)
        ok = validate_paths([cc for cc, _, _ in valid_items], [r.path if r.solved else None for r in results]) \
            if validate else np.ones(len(results), bool)
        for (cc, s, g), r, ns, good in zip(valid_items, results, host_ns, ok):
            if not r.solved:
                res.failed += 1
                continue
            if not good:
                res.path_invalid += 1
            res.rows.append({"set": set_name, "planning_time_us": ns / 1e3,
                             "kernel_time_us": 1e3 * float(r.kernel_time_ms or 0.0), "iterations": int(r.iterations),
                             "cost": float(r.cost), "path_valid": bool(good)})
        if log:
            log(f"{set_name}: {sum(1 for x in res.rows if x['set'] == set_name)} / {len(valid_items)} solved")
    return res


def main(argv=None) -> None:
    import argparse

    ap = argparse.ArgumentParser(description="Evaluate the planner on a problem suite (evaluate_mbm.py style)")
    ap.add_argument("urdf")
    ap.add_argument("problems", help="JSON file: {set: [{valid, start, goals, obstacles}, ...]}")
    ap.add_argument("--srdf")
    ap.add_argument("--batched", action="store_true")
    ap.add_argument("--max-problems-per-set", type=int, default=0)
    ap.add_argument("--max-iterations", type=int, default=5000)
    ap.add_argument("--num-new-samples", type=int, default=64)
    ap.add_argument("--max-nodes", type=int, default=1_000_000)
    ap.add_argument("--step-size", type=float, default=0.5)
    args = ap.parse_args(argv)
    with open(args.urdf) as f:
        urdf = f.read()
    srdf = open(args.srdf).read() if args.srdf else None
    with open(args.problems) as f:
        problems = json.load(f)
    res = evaluate_suite(urdf, srdf, problems, batched=args.batched, max_problems_per_set=args.max_problems_per_set,
                         max_iterations=args.max_iterations, num_new_samples=args.num_new_samples,
                         max_nodes=args.max_nodes, step_size=args.step_size, log=print)
    st = summary(res)
    print(markdown_table(st, "all"))
    print(f"Solved / valid / total: {res.solved} / {res.valid} / {res.total} (paths failing the FP64 re-check: "
          f"{res.path_invalid})")


if __name__ == "__main__":
    main()
