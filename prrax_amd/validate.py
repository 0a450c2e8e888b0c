"""Host-side FP64 validation of returned paths.

The reference validates a planner's paths with an independent signed-distance checker (pyroffi's
robot-world and self distances, ``cuda-rrtc/jax/test_prrtc.py:72-93``; minimum margin > 0 at
``test_prrtc.py:92``).  This module is that checker for the ``collision_context`` wire format,
in float64 and vectorised over configurations:

* FK of the context's joints (``T_j = T_parent(j) . P_j . exp(twist_j q_j)``, twist = (v, w),
  mimic joints ``q = q[mimic] * mul + off``, ``fk_topo_inv`` maps link indices to joints),
* signed distances of every robot sphere to world spheres, capsules, oriented boxes and
  halfspaces (free side +n) and of every self pair (``README.md:77-104``, ``utils.py:58-116``),
* a path is valid when every point ``a + (k / resolution)(b - a)`` of every edge has a minimum
  distance above ``-tol`` (points within ``tol`` of contact cannot be decided in float32).

It does not plan and is not a fallback of any planner call: it only reads paths.
"""

from __future__ import annotations

from typing import Sequence

import numpy as np


def _quat_mats(q: np.ndarray) -> np.ndarray:
    q = q / np.linalg.norm(q, axis=-1, keepdims=True)
    w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1),
    ], -2)


def _exp_se3(twist: np.ndarray, theta: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Rotation [n, 3, 3] and translation [n, 3] of exp(twist * theta) for a batch of angles."""
    v = twist[:3][None, :] * theta[:, None]
    w = twist[3:][None, :] * theta[:, None]
    a = np.linalg.norm(w, axis=1)
    W = np.zeros((theta.shape[0], 3, 3))
    W[:, 0, 1], W[:, 0, 2], W[:, 1, 2] = -w[:, 2], w[:, 1], -w[:, 0]
    W[:, 1, 0], W[:, 2, 0], W[:, 2, 1] = w[:, 2], -w[:, 1], w[:, 0]
    W2 = W @ W
    small = a < 1e-9
    s = np.where(small, 1.0, a)
    A = np.where(small, 1.0, np.sin(s) / s)
    B = np.where(small, 0.5, (1.0 - np.cos(s)) / (s * s))
    C = np.where(small, 1.0 / 6.0, (s - np.sin(s)) / (s * s * s))
    eye = np.eye(3)[None]
    R = eye + A[:, None, None] * W + B[:, None, None] * W2
    V = eye + B[:, None, None] * W + C[:, None, None] * W2
    return R, np.einsum("nij,nj->ni", V, v)


def sphere_centres(cc: dict, qs: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """World centres [n, n_valid, 3] and radii [n_valid] of the robot spheres on valid links."""
    qs = np.asarray(qs, np.float64).reshape(qs.shape[0], -1)
    n = qs.shape[0]
    tw = np.asarray(cc["fk_twists"], np.float64).reshape(-1, 6)
    ptf = np.asarray(cc["fk_parent_tf"], np.float64).reshape(-1, 7)
    par = np.asarray(cc["fk_parent_idx"]).reshape(-1)
    act = np.asarray(cc["fk_act_idx"]).reshape(-1)
    mim = np.asarray(cc["fk_mimic_act_idx"]).reshape(-1)
    mul = np.asarray(cc["fk_mimic_mul"], np.float64).reshape(-1)
    off = np.asarray(cc["fk_mimic_off"], np.float64).reshape(-1)
    inv = np.asarray(cc["fk_topo_inv"]).reshape(-1)
    nj = par.shape[0]
    PR = _quat_mats(ptf[:, :4])
    Rs = np.zeros((nj, n, 3, 3))
    ts = np.zeros((nj, n, 3))
    for j in range(nj):
        if mim[j] >= 0:
            th = qs[:, mim[j]] * mul[j] + off[j]
        elif act[j] >= 0:
            th = qs[:, act[j]]
        else:
            th = np.zeros(n)
        Re, te = _exp_se3(tw[j], th)
        if par[j] >= 0:
            Rb = Rs[par[j]] @ PR[j]
            tb = np.einsum("nij,j->ni", Rs[par[j]], ptf[j, 4:]) + ts[par[j]]
        else:
            Rb = np.broadcast_to(PR[j], (n, 3, 3))
            tb = np.broadcast_to(ptf[j, 4:], (n, 3))
        Rs[j] = Rb @ Re
        ts[j] = np.einsum("nij,nj->ni", Rb, te) + tb
    links = np.asarray(cc["sphere_link_idx"]).reshape(-1)
    loc = np.asarray(cc["sphere_local"], np.float64).reshape(-1, 3)
    rad = np.asarray(cc["sphere_radius"], np.float64).reshape(-1)
    ok = (links >= 0) & (links < nj)
    js = inv[links[ok]]
    P = np.einsum("snij,sj->nsi", Rs[js], loc[ok]) + ts[js].transpose(1, 0, 2)
    return P, rad[ok]


def min_distance(cc: dict, qs, cap: float | None = None) -> np.ndarray:
    """Smallest signed distance [n] over every (robot sphere, obstacle) pair and self pair
    (+inf when there is none).  With ``cap``, a result above ``cap`` is reported as ``cap``:
    sphere pairs farther apart than radii + cap are screened out by their squared distance and
    never take a square root (the validation only needs the sign near contact)."""
    qs = np.asarray(qs, np.float64)
    qs = qs.reshape(qs.shape[0], -1)
    links = np.asarray(cc["sphere_link_idx"]).reshape(-1)
    nj = np.asarray(cc["fk_parent_idx"]).reshape(-1).shape[0]
    best = np.full(qs.shape[0], np.inf if cap is None else float(cap))
    if links.shape[0] == 0 or nj == 0:
        return best
    P, r = sphere_centres(cc, qs)
    n, ns = P.shape[0], P.shape[1]
    flat = P.reshape(-1, 3)
    ws = np.asarray(cc.get("world_spheres", np.zeros((0, 4))), np.float64).reshape(-1, 4)
    if ws.shape[0]:
        # |p - o|^2 = |p|^2 + |o|^2 - 2 p.o as one matrix product (float64: the cancellation
        # costs about 1e-15 m near contact)
        d2 = (flat * flat).sum(1)[:, None] + (ws[:, :3] * ws[:, :3]).sum(1)[None, :] - 2.0 * (flat @ ws[:, :3].T)
        rsum = np.tile(r, n)[:, None] + ws[None, :, 3]
        if cap is None:
            d = np.sqrt(np.maximum(d2, 0.0)) - rsum
            best = np.minimum(best, d.min(axis=1).reshape(n, ns).min(axis=1))
        else:
            lim = np.maximum(rsum + cap, 0.0)
            row, col = np.nonzero(d2 < lim * lim)
            if row.size:
                d = np.sqrt(np.maximum(d2[row, col], 0.0)) - rsum[row, col]
                np.minimum.at(best, row // ns, d)
    wc = np.asarray(cc.get("world_capsules", np.zeros((0, 7))), np.float64).reshape(-1, 7)
    for c in wc:
        a, ab = c[:3], c[3:6] - c[:3]
        rel = flat - a
        t = np.clip((rel @ ab) / max(float(ab @ ab), 1e-300), 0.0, 1.0)
        d = np.linalg.norm(rel - t[:, None] * ab, axis=-1).reshape(n, ns) - r[None] - c[6]
        best = np.minimum(best, d.min(axis=1))
    wb = np.asarray(cc.get("world_boxes", np.zeros((0, 15))), np.float64).reshape(-1, 15)
    for b in wb:
        axes = np.stack([b[3:6], b[6:9], b[9:12]], axis=1)  # columns are the box axes
        qd = np.abs((flat - b[:3]) @ axes) - b[12:15]
        d = (np.sqrt((np.maximum(qd, 0.0) ** 2).sum(-1)) + np.minimum(qd.max(-1), 0.0)).reshape(n, ns) - r[None]
        best = np.minimum(best, d.min(axis=1))
    wh = np.asarray(cc.get("world_halfspaces", np.zeros((0, 6))), np.float64).reshape(-1, 6)
    if wh.shape[0]:
        off = (wh[:, :3] * wh[:, 3:6]).sum(1)
        d = (flat @ wh[:, :3].T - off[None]).reshape(n, ns, -1) - r[None, :, None]
        best = np.minimum(best, d.min(axis=(1, 2)))
    sp = np.asarray(cc.get("self_pairs", np.zeros((0, 2))), np.int64).reshape(-1, 2)
    if sp.shape[0]:
        # self pairs index the full sphere list; map them to the valid spheres
        ok = (links >= 0) & (links < nj)
        pos = np.full(links.shape[0], -1)
        pos[ok] = np.arange(int(ok.sum()))
        good = (sp >= 0).all(1) & (sp < links.shape[0]).all(1)
        sp = sp[good]
        i, j = pos[sp[:, 0]], pos[sp[:, 1]]
        keep = (i >= 0) & (j >= 0)
        i, j = i[keep], j[keep]
        if i.size:
            d = np.linalg.norm(P[:, i] - P[:, j], axis=-1) - r[i][None] - r[j][None]
            best = np.minimum(best, d.min(axis=1))
    return best


def self_pair_distances(cc: dict, q) -> np.ndarray:
    """Signed distance of every self pair of the context at one configuration, in the order of
    ``self_pairs`` (+inf for a pair with an index out of range or a sphere on no joint)."""
    sp = np.asarray(cc.get("self_pairs", np.zeros((0, 2))), np.int64).reshape(-1, 2)
    out = np.full(sp.shape[0], np.inf)
    links = np.asarray(cc["sphere_link_idx"]).reshape(-1)
    nj = np.asarray(cc["fk_parent_idx"]).reshape(-1).shape[0]
    if sp.shape[0] == 0 or links.shape[0] == 0 or nj == 0:
        return out
    P, r = sphere_centres(cc, np.asarray(q, np.float64).reshape(1, -1))
    ok = (links >= 0) & (links < nj)
    pos = np.full(links.shape[0], -1)
    pos[ok] = np.arange(int(ok.sum()))
    good = (sp >= 0).all(1) & (sp < links.shape[0]).all(1)
    i = np.where(good, pos[np.clip(sp[:, 0], 0, links.shape[0] - 1)], -1)
    j = np.where(good, pos[np.clip(sp[:, 1], 0, links.shape[0] - 1)], -1)
    keep = (i >= 0) & (j >= 0)
    out[keep] = np.linalg.norm(P[0, i[keep]] - P[0, j[keep]], axis=-1) - r[i[keep]] - r[j[keep]]
    return out


def path_points(path: np.ndarray, resolution: int = 16) -> np.ndarray:
    """Every point a + (k / resolution)(b - a), k = 0 .. resolution, of every edge of a path."""
    p = np.asarray(path, np.float64)
    t = np.arange(resolution + 1, dtype=np.float64) / resolution
    pts = p[:-1, None, :] + t[None, :, None] * (p[1:] - p[:-1])[:, None, :]
    return pts.reshape(-1, p.shape[1]) if p.shape[0] > 1 else p


def path_min_distance(cc: dict, path: np.ndarray, resolution: int = 16, chunk: int = 512,
                      cap: float | None = None) -> float:
    """Smallest signed distance over the path's edge points (capped at ``cap`` if given)."""
    pts = path_points(path, resolution)
    return float(min(min_distance(cc, pts[i:i + chunk], cap).min() for i in range(0, pts.shape[0], chunk)))


def validate_paths(contexts: dict | Sequence[dict], paths: Sequence[np.ndarray | None], resolution: int = 16,
                   tol: float = 1e-5, workers: int = 1) -> np.ndarray:
    """Per path: True if every edge point clears contact by more than -tol (None paths give
    False).  ``contexts`` is one context for all paths or one per path."""
    ctxs = [contexts] * len(paths) if isinstance(contexts, dict) else list(contexts)

    def one(k):
        p = paths[k]
        if p is None or len(p) == 0:
            return False
        return path_min_distance(ctxs[k], p, resolution, cap=0.01) > -tol

    if workers <= 1:
        return np.array([one(k) for k in range(len(paths))], bool)
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=workers) as ex:
        return np.array(list(ex.map(one, range(len(paths)))), bool)
