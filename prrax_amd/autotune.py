"""Autotuner for batched planning: samples per round and launch shape.

The launch shape (threads per work-group = waves per work-group, team cap, pass-1 budget) never
changes a plan.  The number of samples per round (``num_new_samples``, S) does: it is the
reference's own parameter (prrtc.py:258), and a different S gives different (equally valid)
trees, so a candidate S is only accepted if its solve rate on the tuning sample is at least the
reference default's (S = 128).

Tuning runs the real planner on a stratified subset of the batch: every k-th query plus the
queries with the most obstacles (the long ones), and keeps the candidate with the highest
plans/s.
"""

from __future__ import annotations

import itertools
import time
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from .prrtc import prepare_batch

DEFAULT_SAMPLES = 128


@dataclass
class TuneResult:
    num_new_samples: int
    launch: dict
    plans_per_s: float
    solve_rate: float
    table: list = field(default_factory=list)

    def kwargs(self) -> dict:
        """Keyword arguments for ``prepare_batch`` / ``prrtc_plan_batch``."""
        return dict(num_new_samples=self.num_new_samples, **self.launch)


def tuning_sample(contexts: Sequence[dict], n: int = 96, n_heavy: int = 16) -> list[int]:
    """Indices of a stratified subset: ``n_heavy`` queries with the most world primitives plus
    evenly spaced others, ``n`` in total."""
    B = len(contexts)
    if B <= n:
        return list(range(B))
    load = np.array([c["world_spheres"].shape[0] + 4 * c["world_boxes"].shape[0] + 2 * c["world_capsules"].shape[0]
                     for c in contexts])
    heavy = list(np.argsort(-load, kind="stable")[:n_heavy])
    rest = [i for i in np.linspace(0, B - 1, n - n_heavy).astype(int) if i not in heavy]
    return sorted(set(int(i) for i in heavy + rest))[:n]


def autotune(starts, goals, collision_context, *, max_nodes_per_query=None, samples: Sequence[int] = (128,),
             threads: Sequence[int] = (256, 512), pass1_rounds: Sequence[int] = (24,),
             team_max: Sequence[int] = (0,), n_sample: int = 96, repeats: int = 1,
             log=None, **plan_kw) -> TuneResult:
    """Times every candidate on a tuning sample of the batch and returns the fastest whose solve
    rate is at least the default S's.  ``plan_kw`` are the planner keywords of the real run
    (max_iterations, step_size, limits, ...); ``collision_context`` is one dict or one per query.

    By default only the work-group size is tuned (S stays 128, the team cap stays the library's):
    a sample rarely holds the few queries that never connect, and those set a batch's makespan,
    so an S or a team cap chosen on the sample can lose on the whole batch.  ``pass1_rounds`` is pinned (24, the full-batch default) so the
    sample runs in the same regime as a batch larger than the resident work-groups."""
    B = len(starts)
    ctxs = collision_context if isinstance(collision_context, (list, tuple)) else [collision_context] * B
    idx = tuning_sample(ctxs, n_sample)
    st = np.stack([np.asarray(starts[i], np.float32) for i in idx])
    gl = [goals[i] for i in idx]
    cx = [ctxs[i] for i in idx] if isinstance(collision_context, (list, tuple)) else collision_context
    mn = None
    if max_nodes_per_query is not None:
        mn = [int(max_nodes_per_query[i]) for i in idx]
        plan_kw = dict(plan_kw, max_nodes=max(mn))
    table = []
    for S, nt, p1, tm in itertools.product(samples, threads, pass1_rounds, team_max):
        pb = prepare_batch(st, gl, collision_context=cx, max_nodes_per_query=mn, num_new_samples=S, threads=nt,
                           pass1_rounds=p1, team_max=tm, **plan_kw)
        pb.enqueue()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(repeats):
            pb.enqueue()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / repeats
        solve = float(np.mean(pb.bufs.host_summary()["solved"] == 1))
        row = {"num_new_samples": S, "threads": nt, "pass1_rounds": p1, "team_max": tm,
               "plans_per_s": len(idx) / dt, "solve_rate": solve}
        table.append(row)
        if log:
            log(f"autotune: {row}")
        del pb
    base = [r["solve_rate"] for r in table if r["num_new_samples"] == DEFAULT_SAMPLES]
    floor = max(base) if base else 0.0
    ok = [r for r in table if r["solve_rate"] >= floor] or table
    best = max(ok, key=lambda r: r["plans_per_s"])
    return TuneResult(best["num_new_samples"],
                      {"threads": best["threads"], "pass1_rounds": best["pass1_rounds"], "team_max": best["team_max"]},
                      best["plans_per_s"], best["solve_rate"], table)
