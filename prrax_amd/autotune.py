"""Autotuner for batched planning: samples per round, waves per query, and the FK/check tile.

Knobs (every one but S leaves the plans unchanged):

* ``num_new_samples`` (S, the reference's own parameter, prrtc.py:258): explorers per round.  A
  different S gives different (equally valid) trees, so a candidate S is accepted only if its
  solve rate on the tuning sample is at least the reference default's (S = 128).
* ``threads`` (waves per work-group) and ``team_max`` (the largest team of work-groups one open
  query gets): together the waves per query.
* ``fk_batch`` (8 or 16 configurations per wave): the tile of one FK + collision-check step;
  a check work unit covers at most this many waypoints.

The makespan of an uneven batch is set by its few longest queries (those that never connect
run until their trees are full), so the tuning sample must hold them: a short probe launch of
the whole batch (``probe_rounds`` rounds) finds the queries still open after it, and the sample
is those plus an evenly spaced subset of the rest, as many queries as the GPU holds work-groups
(256), so that the one-work-group pass is loaded as in the whole batch.  Each candidate then
plans the sample with the real settings, and the fastest at an accepted solve rate wins.  The search is staged: S
first at the library's launch defaults, then the launch shape at the chosen S.
"""

from __future__ import annotations

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
    sample: list = field(default_factory=list)
    tail: list = field(default_factory=list)

    def kwargs(self) -> dict:
        """Keyword arguments for ``prepare_batch`` / ``prrtc_plan_batch``."""
        return dict(num_new_samples=self.num_new_samples, **self.launch)


def _subset(starts, goals, ctxs, shared_ctx, mnq, idx):
    st = np.stack([np.asarray(starts[i], np.float32) for i in idx])
    gl = [goals[i] for i in idx]
    cx = shared_ctx if shared_ctx is not None else [ctxs[i] for i in idx]
    mn = None if mnq is None else [int(mnq[i]) for i in idx]
    return st, gl, cx, mn


def probe_tail(starts, goals, collision_context, *, max_nodes_per_query=None, probe_rounds: int = 24,
               num_new_samples: int = DEFAULT_SAMPLES, **plan_kw) -> np.ndarray:
    """Indices of the queries still open after ``probe_rounds`` rounds of the whole batch (one
    launch), longest trees first."""
    kw = dict(plan_kw, max_iterations=probe_rounds)
    if max_nodes_per_query is not None:
        kw["max_nodes"] = int(max(max_nodes_per_query))
    pb = prepare_batch(np.stack([np.asarray(s, np.float32) for s in starts]), goals,
                       collision_context=collision_context, max_nodes_per_query=max_nodes_per_query,
                       num_new_samples=num_new_samples, **kw)
    pb.enqueue()
    torch.cuda.synchronize()
    h = pb.bufs.host_summary()
    open_q = np.nonzero(h["solved"] != 1)[0]
    size = h["sizes"].sum(axis=1)
    return open_q[np.argsort(-size[open_q], kind="stable")]


def tuning_sample(n_batch: int, tail: Sequence[int], n: int = 256, n_tail: int = 16) -> list[int]:
    """The first ``n_tail`` tail queries plus evenly spaced others, ``n`` in total."""
    if n_batch <= n:
        return list(range(n_batch))
    head = [int(i) for i in list(tail)[:n_tail]]
    rest = [int(i) for i in np.linspace(0, n_batch - 1, n - len(head)).astype(int) if int(i) not in head]
    return sorted(set(head + rest))[:n]


def autotune(starts, goals, collision_context, *, max_nodes_per_query=None,
             samples: Sequence[int] = (128, 256, 512), threads: Sequence[int] = (256, 512),
             team_max: Sequence[int] = (0, 128), fk_batch: Sequence[int] = (16, 8), n_sample: int = 256,
             n_tail: int = 16, probe_rounds: int = 24, repeats: int = 1, min_gain: float = 0.03, log=None,
             **plan_kw) -> TuneResult:
    """Times candidates on a tail-seeded sample of the batch and returns the fastest whose solve
    rate is at least S = 128's.  A candidate replaces the defaults of its stage (S = 128; then the
    library's launch shape at the chosen S) only if it is faster by more than ``min_gain`` on
    the sample: smaller differences are within what a sample predicts of the whole batch.
    ``plan_kw`` are the planner keywords of the real run (max_iterations, step_size, limits,
    ...); ``collision_context`` is one dict or one per query."""
    B = len(starts)
    shared = None if isinstance(collision_context, (list, tuple)) else collision_context
    ctxs = collision_context if shared is None else None
    tail = probe_tail(starts, goals, collision_context, max_nodes_per_query=max_nodes_per_query,
                      probe_rounds=probe_rounds, **plan_kw)
    idx = tuning_sample(B, tail, n_sample, n_tail)
    st, gl, cx, mn = _subset(starts, goals, ctxs, shared, max_nodes_per_query, idx)
    kw = dict(plan_kw)
    if mn is not None:
        kw["max_nodes"] = max(mn)
    if log:
        log(f"autotune: {len(tail)} queries open after {probe_rounds} rounds; sample of {len(idx)} holds "
            f"{sum(1 for i in idx if i in set(int(t) for t in tail))} of them")
    table: list[dict] = []

    def run(S, nt, tm, fb) -> dict:
        for r in table:
            if (r["num_new_samples"], r["threads"], r["team_max"], r["fk_batch"]) == (S, nt, tm, fb):
                return r
        # pass 1 runs 24 rounds, as for a batch that fills the resident work-groups (a sample
        # smaller than that would otherwise start every query on a team)
        pb = prepare_batch(st, gl, collision_context=cx, max_nodes_per_query=mn, num_new_samples=S, threads=nt,
                           team_max=tm, fk_batch=fb, pass1_rounds=24, **kw)
        pb.enqueue()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(repeats):
            pb.enqueue()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / repeats
        solve = float(np.mean(pb.bufs.host_summary()["solved"] == 1))
        row = {"num_new_samples": S, "threads": nt, "team_max": tm, "fk_batch": fb,
               "plans_per_s": len(idx) / dt, "ms": 1000.0 * dt, "solve_rate": solve}
        table.append(row)
        if log:
            log(f"autotune: {row}")
        del pb
        return row

    # stage 1: samples per round at the library's launch defaults
    base = {S: run(S, 0, 0, 0) for S in samples}
    floor = base[DEFAULT_SAMPLES]["solve_rate"] if DEFAULT_SAMPLES in base else 0.0
    ok = [r for r in base.values() if r["solve_rate"] >= floor] or list(base.values())
    top = max(ok, key=lambda r: r["plans_per_s"])
    dflt = base.get(DEFAULT_SAMPLES, top)
    S_best = top["num_new_samples"] if top["plans_per_s"] > (1.0 + min_gain) * dflt["plans_per_s"] \
        else dflt["num_new_samples"]
    # stage 2: the launch shape at that S (results do not depend on it)
    for nt in threads:
        for tm in team_max:
            for fb in fk_batch:
                run(S_best, nt, tm, fb)
    cands = [r for r in table if r["num_new_samples"] == S_best]
    best = max(cands, key=lambda r: r["plans_per_s"])
    shape0 = base[S_best]  # the library's launch defaults at S_best
    if best["plans_per_s"] <= (1.0 + min_gain) * shape0["plans_per_s"]:
        best = shape0
    return TuneResult(best["num_new_samples"],
                      {"threads": best["threads"], "team_max": best["team_max"], "fk_batch": best["fk_batch"]},
                      best["plans_per_s"], best["solve_rate"], table, idx, [int(t) for t in tail])
