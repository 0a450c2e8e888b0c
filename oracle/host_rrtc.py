"""The reference's CPU planner path, restated in NumPy.  TEST INFRASTRUCTURE ONLY.

Only ``bench.py``'s ``cpu_baseline`` leg and ``tests/`` may import this module.

The reference drives batched RRT-Connect from the host and validates each round's candidate
edges with one batched FK + collision call (``cuda-rrtc/jax/prrtc_robogpu.py:183-264``); with
pyroffi's JAX-on-CPU checker that is the reference's CPU path.  Neither pyroffi nor JAX exists
here, so this module restates both halves:

* ``plan`` follows ``prrtc_robogpu.py:183-264`` step for step: uniform samples from
  ``default_rng(seed)``, the active tree's nearest nodes (``np.argmin``), steer by at most
  ``step_size`` (``prrtc_robogpu.py:36-46``), one batched check of all extension edges at
  ``linspace(0, 1, G)`` (``prrtc_robogpu.py:49-60``), inserts, then a greedy connect of the other
  tree toward every new node for up to ``connect_max_steps`` steps (a step that lands within
  ``step_size * 1e-3 + 1e-6`` of its target connects the trees), and the trees swap roles every
  iteration.  Like the reference, the goal tree is rooted at the first goal only.
* ``BatchChecker`` is the batched FK + collision predicate in float32 NumPy (the JAX-on-CPU
  analogue): FK by twist exponentials through the parent chain, then "any pair distance <= 0"
  against world spheres, capsules, boxes and halfspaces and between self pairs.
"""

from __future__ import annotations

import numpy as np

F32 = np.float32


class BatchChecker:
    """Vectorised float32 FK + binary collision check of configurations ``[n, dim]``."""

    def __init__(self, cc: dict, dim: int):
        self.dim = dim
        g = lambda k, f, dt=F32: np.asarray(cc.get(k, np.zeros((0, f), dt)), dt).reshape(-1, f)  # noqa: E731
        self.tw = g("fk_twists", 6)
        self.ptf = g("fk_parent_tf", 7)
        self.par = np.asarray(cc["fk_parent_idx"]).reshape(-1)
        self.act = np.asarray(cc["fk_act_idx"]).reshape(-1)
        self.mim = np.asarray(cc["fk_mimic_act_idx"]).reshape(-1)
        self.mul = np.asarray(cc["fk_mimic_mul"], F32).reshape(-1)
        self.off = np.asarray(cc["fk_mimic_off"], F32).reshape(-1)
        inv = np.asarray(cc["fk_topo_inv"]).reshape(-1)
        nj = self.par.shape[0]
        links = np.asarray(cc["sphere_link_idx"]).reshape(-1)
        ok = (links >= 0) & (links < nj)
        self.sj = inv[links[ok]]
        self.loc = np.asarray(cc["sphere_local"], F32).reshape(-1, 3)[ok]
        self.rad = np.asarray(cc["sphere_radius"], F32).reshape(-1)[ok]
        pos = np.full(links.shape[0], -1)
        pos[ok] = np.arange(int(ok.sum()))
        sp = g("self_pairs", 2, np.int64)
        sp = sp[(sp >= 0).all(1) & (sp < links.shape[0]).all(1)] if sp.size else sp
        i, j = (pos[sp[:, 0]], pos[sp[:, 1]]) if sp.size else (np.zeros(0, int), np.zeros(0, int))
        keep = (i >= 0) & (j >= 0)
        self.pi, self.pj = i[keep], j[keep]
        self.ws, self.wc, self.wb, self.wh = g("world_spheres", 4), g("world_capsules", 7), g("world_boxes", 15), \
            g("world_halfspaces", 6)
        q = self.ptf[:, :4] / np.linalg.norm(self.ptf[:, :4], axis=1, keepdims=True)
        w, x, y, z = q.T
        self.PR = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
                            np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
                            np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)],
                           -2).astype(F32)
        self.enabled = nj > 0 and self.rad.shape[0] > 0 and (
            self.ws.shape[0] + self.wc.shape[0] + self.wb.shape[0] + self.wh.shape[0] + self.pi.shape[0]) > 0

    def _centres(self, qs: np.ndarray) -> np.ndarray:
        n = qs.shape[0]
        nj = self.par.shape[0]
        Rs = np.zeros((nj, n, 3, 3), F32)
        ts = np.zeros((nj, n, 3), F32)
        for j in range(nj):
            if self.mim[j] >= 0:
                th = qs[:, self.mim[j]] * self.mul[j] + self.off[j]
            elif self.act[j] >= 0:
                th = qs[:, self.act[j]]
            else:
                th = np.zeros(n, F32)
            v = self.tw[j, :3][None] * th[:, None]
            w = self.tw[j, 3:][None] * th[:, None]
            a = np.linalg.norm(w, axis=1)
            W = np.zeros((n, 3, 3), F32)
            W[:, 0, 1], W[:, 0, 2], W[:, 1, 2] = -w[:, 2], w[:, 1], -w[:, 0]
            W[:, 1, 0], W[:, 2, 0], W[:, 2, 1] = w[:, 2], -w[:, 1], w[:, 0]
            W2 = W @ W
            s = np.where(a < 1e-6, F32(1), a)
            A = np.where(a < 1e-6, F32(1), np.sin(s) / s)[:, None, None]
            B = np.where(a < 1e-6, F32(0.5), (1 - np.cos(s)) / (s * s))[:, None, None]
            C = np.where(a < 1e-6, F32(1 / 6), (s - np.sin(s)) / (s * s * s))[:, None, None]
            eye = np.eye(3, dtype=F32)[None]
            Re = eye + A * W + B * W2
            te = np.einsum("nij,nj->ni", eye + B * W + C * W2, v)
            if self.par[j] >= 0:
                Rb = Rs[self.par[j]] @ self.PR[j]
                tb = Rs[self.par[j]] @ self.ptf[j, 4:] + ts[self.par[j]]
            else:
                Rb = np.broadcast_to(self.PR[j], (n, 3, 3))
                tb = np.broadcast_to(self.ptf[j, 4:], (n, 3))
            Rs[j] = Rb @ Re
            ts[j] = np.einsum("nij,nj->ni", Rb, te) + tb
        return np.einsum("snij,sj->nsi", Rs[self.sj], self.loc) + ts[self.sj].transpose(1, 0, 2)

    def free(self, qs: np.ndarray) -> np.ndarray:
        """True for each configuration with no pair distance <= 0."""
        qs = np.asarray(qs, F32).reshape(-1, self.dim)
        n = qs.shape[0]
        if not self.enabled or n == 0:
            return np.ones(n, bool)
        return ~self._hits(self._centres(qs))

    def _hits(self, P: np.ndarray) -> np.ndarray:
        """Contact of the sphere centres P [n, spheres, 3] with the world or a self pair."""
        n = P.shape[0]
        hit = np.zeros(n, bool)
        r = self.rad
        if self.ws.shape[0]:
            # |p - o|^2 = |p|^2 + |o|^2 - 2 p.o: one matrix product, evaluated in float64 so that
            # the cancellation stays far below float32 rounding
            flat = P.reshape(-1, 3).astype(np.float64)
            o = self.ws[:, :3].astype(np.float64)
            d2 = (flat * flat).sum(1)[:, None] + (o * o).sum(1)[None] - 2.0 * (flat @ o.T)
            rr = (np.tile(r, n)[:, None] + self.ws[None, :, 3]).astype(np.float64)
            hit |= ((rr >= 0) & (d2 <= rr * rr)).reshape(n, -1).any(axis=1)
        for c in self.wc:
            a, ab = c[:3], c[3:6] - c[:3]
            rel = P - a
            t = np.clip((rel @ ab) / max(float(ab @ ab), 1e-30), 0, 1)
            d2 = ((rel - t[..., None] * ab) ** 2).sum(-1)
            rr = r[None] + c[6]
            hit |= ((rr >= 0) & (d2 <= rr * rr)).any(axis=1)
        for b in self.wb:
            qd = np.abs((P - b[:3]) @ np.stack([b[3:6], b[6:9], b[9:12]], 1)) - b[12:15]
            d2 = (np.maximum(qd, 0) ** 2).sum(-1)
            hit |= np.where(r[None] >= 0, d2 <= r[None] * r[None], qd.max(-1) <= r[None]).any(axis=1)
        for h in self.wh:
            hit |= ((P @ h[:3] - F32(h[:3] @ h[3:6])) <= r[None]).any(axis=1)
        if self.pi.shape[0]:
            d2 = ((P[:, self.pi] - P[:, self.pj]) ** 2).sum(-1)
            rr = r[self.pi] + r[self.pj]
            hit |= ((rr >= 0)[None] & (d2 <= (rr * rr)[None])).any(axis=1)
        return hit


class CloudChecker(BatchChecker):
    """BatchChecker plus an environment point cloud, each point a sphere of radius ``r_env``
    (the point-cloud planner's checker, prrtc_robogpu.py:103-298): a k-d tree over the points
    (scipy) gives every robot sphere's nearest point in float64, and a sphere within
    r + r_env of it is in contact."""

    def __init__(self, cc: dict, dim: int, points, r_env: float):
        super().__init__(cc, dim)
        from scipy.spatial import cKDTree

        pts = np.asarray(points, np.float64).reshape(-1, 3)
        self.tree = cKDTree(pts) if pts.shape[0] else None
        self.r_env = float(r_env)

    def free(self, qs: np.ndarray) -> np.ndarray:
        qs = np.asarray(qs, F32).reshape(-1, self.dim)
        n = qs.shape[0]
        if n == 0 or self.par.shape[0] == 0 or self.rad.shape[0] == 0:
            return np.ones(n, bool)
        P = self._centres(qs)
        hit = self._hits(P) if self.enabled else np.zeros(n, bool)
        if self.tree is not None:
            reach = self.rad.astype(np.float64) + self.r_env
            d, _ = self.tree.query(P.reshape(-1, 3).astype(np.float64), k=1,
                                   distance_upper_bound=float(reach.max()) * (1 + 1e-9) + 1e-12)
            hit |= (d.reshape(n, -1) <= reach[None]).any(axis=1)
        return ~hit


def _steer(q_from, q_to, step):
    delta = q_to - q_from
    dist = np.linalg.norm(delta, axis=1, keepdims=True)
    return q_from + delta * np.minimum(1.0, step / np.maximum(dist, 1e-8))


class _Tree:
    def __init__(self, root, max_nodes, dim):
        self.cfg = np.zeros((max_nodes, dim), F32)
        self.par = np.full(max_nodes, -1, np.int32)
        self.cfg[0], self.par[0], self.size = root, 0, 1

    def add(self, q, parent) -> int:
        if self.size >= self.cfg.shape[0]:
            return -1
        i = self.size
        self.cfg[i], self.par[i] = q, parent
        self.size += 1
        return i

    def nearest(self, qs):
        v = self.cfg[:self.size]
        d2 = np.einsum("ksd,ksd->ks", v[:, None, :] - qs[None], v[:, None, :] - qs[None])
        return np.argmin(d2, axis=0)

    def walk(self, i):
        out = []
        for _ in range(self.size):
            out.append(self.cfg[i])
            if self.par[i] == i:
                break
            i = int(self.par[i])
        return out[::-1]


def plan(checker: BatchChecker, start, goals, *, max_iterations=10_000, step_size=0.5, num_new_samples=128,
         granularity=16, max_nodes=16384, connect_max_steps=32, min_vals=None, max_vals=None, seed=0,
         time_limit: float | None = None) -> dict:
    """One query; returns solved, path (start -> goal), iterations, tree sizes and whether the
    optional wall-time limit (seconds) ended it."""
    import time

    t_end = None if time_limit is None else time.perf_counter() + time_limit
    dim = checker.dim
    start = np.asarray(start, F32).reshape(dim)
    goal = np.asarray(goals, F32).reshape(-1, dim)[0]
    lo = np.full(dim, -np.pi, F32) if min_vals is None else np.asarray(min_vals, F32)
    hi = np.full(dim, np.pi, F32) if max_vals is None else np.asarray(max_vals, F32)
    rng = np.random.default_rng(seed)
    ta, tb = _Tree(start, max_nodes, dim), _Tree(goal, max_nodes, dim)
    ts = np.linspace(0.0, 1.0, granularity, dtype=F32)

    def edges_free(qa, qb):
        wp = qa[:, None, :] * (1 - ts)[None, :, None] + qb[:, None, :] * ts[None, :, None]
        return checker.free(wp.reshape(-1, dim)).reshape(qa.shape[0], granularity).all(axis=1)

    active, other = ta, tb
    conn, it, timed_out = None, 0, False
    for it in range(1, max_iterations + 1):
        if t_end is not None and time.perf_counter() > t_end:
            timed_out = True
            break
        smp = rng.uniform(lo, hi, size=(num_new_samples, dim)).astype(F32)
        near = active.nearest(smp)
        qn = active.cfg[near]
        qnew = _steer(qn, smp, step_size).astype(F32)
        ok = edges_free(qn, qnew)
        new_ids, new_q = [], []
        for s in np.nonzero(ok)[0]:
            i = active.add(qnew[s], int(near[s]))
            if i < 0:
                break
            new_ids.append(i)
            new_q.append(qnew[s])
        if not new_ids:
            active, other = other, active
            continue
        tgt = np.stack(new_q)
        oidx = other.nearest(tgt)
        live = np.ones(tgt.shape[0], bool)
        for _ in range(connect_max_steps):
            if not live.any() or conn is not None:
                break
            li = np.nonzero(live)[0]
            qf = other.cfg[oidx[li]]
            qs = _steer(qf, tgt[li], step_size).astype(F32)
            fc = edges_free(qf, qs)
            for k, s in enumerate(li):
                if not fc[k]:
                    live[s] = False
                    continue
                i = other.add(qs[k], int(oidx[s]))
                if i < 0:
                    live[s] = False
                    continue
                oidx[s] = i
                if np.linalg.norm(qs[k] - tgt[s]) <= step_size * 1e-3 + 1e-6:
                    conn = (new_ids[s], i) if active is ta else (i, new_ids[s])
                    break
        if conn is not None:
            break
        active, other = other, active
    path = None
    if conn is not None:
        pb = tb.walk(conn[1])
        path = np.stack(ta.walk(conn[0]) + pb[::-1])
    return {"solved": conn is not None, "path": path, "iterations": it, "sizes": (ta.size, tb.size),
            "timed_out": timed_out}
