"""Python face of the CPU oracle.  TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module; the product package never does.

Two halves:

* ``liboracle.so`` (``prrtc_oracle.c``): the FP32 restatement of the reference planner kernel
  (``cuda-rrtc/CUDA/prrtc_planner.cu:675-1066``) under the lockstep schedule, with the FK,
  collision predicates, Halton sampler and nearest-neighbour search it uses.  The GPU library
  must match it bit for bit.
* ``*64`` functions below: an FP64 numpy restatement written independently from the
  mathematics (4x4 homogeneous transforms, Rodrigues' formula, closed-form point/segment/box/
  plane distances).  These pin the FP32 oracle: FK poses within 1e-5 m/rad and collision
  predicates equal away from contact.
"""

from __future__ import annotations

import ctypes
import math
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int)

STATUS_SOLVED = 1
STATUS_ITER_CAP = 2
STATUS_TREE_FULL = 3


class OcContext(ctypes.Structure):
    _fields_ = [
        ("n_joints", ctypes.c_int), ("dim", ctypes.c_int),
        ("twists", _f32p), ("parent_tf", _f32p), ("parent_idx", _i32p), ("act_idx", _i32p),
        ("mimic_mul", _f32p), ("mimic_off", _f32p), ("mimic_act_idx", _i32p), ("topo_inv", _i32p),
        ("n_spheres", ctypes.c_int), ("sphere_link_idx", _i32p), ("sphere_local", _f32p),
        ("sphere_radius", _f32p),
        ("n_world_spheres", ctypes.c_int), ("world_spheres", _f32p),
        ("n_world_capsules", ctypes.c_int), ("world_capsules", _f32p),
        ("n_world_boxes", ctypes.c_int), ("world_boxes", _f32p),
        ("n_world_halfspaces", ctypes.c_int), ("world_halfspaces", _f32p),
        ("n_self_pairs", ctypes.c_int), ("self_pairs", _i32p),
        ("collision_mode", ctypes.c_int), ("collision_margin", ctypes.c_float),
    ]


class OcParams(ctypes.Structure):
    _fields_ = [
        ("max_iterations", ctypes.c_int), ("step_size", ctypes.c_float),
        ("num_new_samples", ctypes.c_int), ("balance_mode", ctypes.c_int),
        ("tree_ratio", ctypes.c_float), ("dynamic_domain", ctypes.c_int),
        ("dd_alpha", ctypes.c_float), ("dd_radius", ctypes.c_float), ("dd_min_radius", ctypes.c_float),
        ("dim", ctypes.c_int), ("max_nodes", ctypes.c_int), ("granularity", ctypes.c_int),
        ("seed", ctypes.c_uint),
    ]


class OcOutputs(ctypes.Structure):
    _fields_ = [
        ("tree_a", _f32p), ("tree_b", _f32p), ("parents_a", _i32p), ("parents_b", _i32p),
        ("radii_a", _f32p), ("radii_b", _f32p),
        ("sizes", ctypes.c_int * 2), ("iterations", ctypes.c_int), ("conn", ctypes.c_int * 3),
        ("status", ctypes.c_int), ("fkcc_evals", ctypes.c_longlong),
    ]


_LIB = None


def bind(L: ctypes.CDLL) -> ctypes.CDLL:
    """Declares the oracle's C signatures on a loaded library (this one, or a variant of it
    built with a planted mistake)."""
    L.oc_fk.argtypes = [ctypes.POINTER(OcContext), _f32p, _f32p]
    L.oc_in_collision.argtypes = [ctypes.POINTER(OcContext), _f32p]
    L.oc_in_collision.restype = ctypes.c_int
    L.oc_in_collision_batch.argtypes = [ctypes.POINTER(OcContext), _f32p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint8)]
    L.oc_min_distance_batch.argtypes = [ctypes.POINTER(OcContext), _f32p, ctypes.c_int, _f32p]
    L.oc_sphere_positions.argtypes = [ctypes.POINTER(OcContext), _f32p, _f32p]
    L.oc_cloud_hit_batch.argtypes = [ctypes.POINTER(OcContext), _f32p, ctypes.c_int, _f32p, ctypes.c_int,
                                     ctypes.c_float, ctypes.POINTER(ctypes.c_uint8)]
    L.oc_halton.argtypes = [ctypes.c_int, ctypes.c_int]
    L.oc_halton.restype = ctypes.c_float
    L.oc_prime_order.argtypes = [ctypes.c_int, ctypes.c_uint, _i32p]
    L.oc_sincos.argtypes = [ctypes.c_float, _f32p, _f32p]
    L.oc_nearest.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_int, _i32p, _f32p]
    L.oc_dd_shrink.argtypes = [_f32p, ctypes.c_float, ctypes.c_float, ctypes.c_float]
    L.oc_dd_grow.argtypes = [_f32p, ctypes.c_float]
    L.oc_dd_skips.argtypes = [ctypes.c_float, ctypes.c_float]
    L.oc_dd_skips.restype = ctypes.c_int
    L.oc_balance_side.argtypes = [ctypes.c_int] * 6 + [ctypes.c_float, _i32p]
    L.oc_balance_side.restype = ctypes.c_int
    L.oc_steer.argtypes = [_f32p, _f32p, ctypes.c_int, ctypes.c_float, ctypes.c_float, _f32p]
    L.oc_connect_steps.argtypes = [_f32p, _f32p, ctypes.c_int, ctypes.c_float, ctypes.c_float, _f32p]
    L.oc_connect_steps.restype = ctypes.c_int
    L.oc_plan.argtypes = [ctypes.POINTER(OcContext), ctypes.POINTER(OcParams), _f32p, _f32p, ctypes.c_int,
                          _f32p, _f32p, ctypes.POINTER(OcOutputs)]
    L.oc_plan.restype = ctypes.c_int
    L.oc_trace_path.argtypes = [_f32p, _f32p, _i32p, _i32p, ctypes.c_int, ctypes.c_int, _i32p, _f32p, ctypes.c_int]
    L.oc_trace_path.restype = ctypes.c_int
    return L


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            from oracle import build as _b  # noqa: PLC0415

            _b.build()
        _LIB = bind(ctypes.CDLL(str(LIB_PATH)))
    return _LIB


def _fp(a):
    return a.ctypes.data_as(_f32p)


def _ip(a):
    return a.ctypes.data_as(_i32p)


_ROBOT_KEYS = ("fk_twists", "fk_parent_tf", "fk_parent_idx", "fk_act_idx", "fk_mimic_mul", "fk_mimic_off",
               "fk_mimic_act_idx", "fk_topo_inv", "sphere_link_idx", "sphere_local", "sphere_radius")
_INT_KEYS = {"fk_parent_idx", "fk_act_idx", "fk_mimic_act_idx", "fk_topo_inv", "sphere_link_idx", "self_pairs"}
_WORLD = (("world_spheres", 4), ("world_capsules", 7), ("world_boxes", 15), ("world_halfspaces", 6))


class Context:
    """Owns contiguous copies of a collision_context and the ``oc_context`` pointing at them."""

    def __init__(self, cc: dict, dim: int, collision_mode: int = 0, collision_margin: float = 0.02):
        self.arrays = {}
        for k in _ROBOT_KEYS:
            dt = np.int32 if k in _INT_KEYS else np.float32
            self.arrays[k] = np.ascontiguousarray(np.asarray(cc[k], dtype=dt))
        for k, f in _WORLD:
            a = cc.get(k, np.zeros((0, f), np.float32))
            self.arrays[k] = np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1, f))
        self.arrays["self_pairs"] = np.ascontiguousarray(
            np.asarray(cc.get("self_pairs", np.zeros((0, 2), np.int32)), np.int32).reshape(-1, 2))
        a = self.arrays
        self.dim = dim
        self.n_joints = int(a["fk_parent_idx"].shape[0])
        self.c = OcContext(
            self.n_joints, dim, _fp(a["fk_twists"]), _fp(a["fk_parent_tf"]), _ip(a["fk_parent_idx"]),
            _ip(a["fk_act_idx"]), _fp(a["fk_mimic_mul"]), _fp(a["fk_mimic_off"]), _ip(a["fk_mimic_act_idx"]),
            _ip(a["fk_topo_inv"]), int(a["sphere_link_idx"].shape[0]), _ip(a["sphere_link_idx"]),
            _fp(a["sphere_local"]), _fp(a["sphere_radius"]),
            a["world_spheres"].shape[0], _fp(a["world_spheres"]), a["world_capsules"].shape[0],
            _fp(a["world_capsules"]), a["world_boxes"].shape[0], _fp(a["world_boxes"]),
            a["world_halfspaces"].shape[0], _fp(a["world_halfspaces"]),
            a["self_pairs"].shape[0], _ip(a["self_pairs"]), int(collision_mode), float(collision_margin))


# --------------------------------------------------------------------------- FP32 oracle
def sincos(x: float) -> tuple[float, float]:
    s, c = ctypes.c_float(), ctypes.c_float()
    lib().oc_sincos(ctypes.c_float(x), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def fk(ctx: Context, q) -> np.ndarray:
    q = np.ascontiguousarray(np.asarray(q, np.float32).reshape(-1))
    out = np.zeros((ctx.n_joints, 7), np.float32)
    lib().oc_fk(ctypes.byref(ctx.c), _fp(q), _fp(out))
    return out


def sphere_positions(ctx: Context, q) -> np.ndarray:
    q = np.ascontiguousarray(np.asarray(q, np.float32).reshape(-1))
    out = np.zeros((ctx.arrays["sphere_radius"].shape[0], 3), np.float32)
    lib().oc_sphere_positions(ctypes.byref(ctx.c), _fp(q), _fp(out))
    return out


def in_collision(ctx: Context, qs, L=None) -> np.ndarray:
    qs = np.ascontiguousarray(np.asarray(qs, np.float32).reshape(-1, ctx.dim))
    out = np.zeros(qs.shape[0], np.uint8)
    (L or lib()).oc_in_collision_batch(ctypes.byref(ctx.c), _fp(qs), qs.shape[0], out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return out.astype(bool)


def cloud_hits(ctx: Context, qs, points, r_env: float) -> np.ndarray:
    """Per configuration: some robot sphere within r + r_env of a cloud point (brute force)."""
    qs = np.ascontiguousarray(np.asarray(qs, np.float32).reshape(-1, ctx.dim))
    pts = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 3))
    out = np.zeros(qs.shape[0], np.uint8)
    lib().oc_cloud_hit_batch(ctypes.byref(ctx.c), _fp(qs), qs.shape[0], _fp(pts), pts.shape[0], float(r_env),
                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return out.astype(bool)


def min_distance(ctx: Context, qs) -> np.ndarray:
    qs = np.ascontiguousarray(np.asarray(qs, np.float32).reshape(-1, ctx.dim))
    out = np.zeros(qs.shape[0], np.float32)
    lib().oc_min_distance_batch(ctypes.byref(ctx.c), _fp(qs), qs.shape[0], _fp(out))
    return out


def halton(base: int, index: int) -> float:
    return lib().oc_halton(int(base), int(index))


def prime_order(explorer: int, seed: int = 0) -> np.ndarray:
    out = np.zeros(16, np.int32)
    lib().oc_prime_order(int(explorer), ctypes.c_uint(seed), _ip(out))
    return out


def nearest(tree: np.ndarray, size: int, queries) -> tuple[np.ndarray, np.ndarray]:
    """Lowest-index argmin of squared distance; ``tree`` is SoA ``[dim, max_nodes]``."""
    tree = np.ascontiguousarray(tree, np.float32)
    dim, max_nodes = tree.shape
    qs = np.ascontiguousarray(np.asarray(queries, np.float32).reshape(-1, dim))
    idx = np.zeros(qs.shape[0], np.int32)
    dsq = np.zeros(qs.shape[0], np.float32)
    lib().oc_nearest(_fp(tree), max_nodes, int(size), dim, _fp(qs), qs.shape[0], _ip(idx), _fp(dsq))
    return idx, dsq


# ---- planner rules (prrtc_planner.cu:46-83, 790-824, 864-876, 975-993); L: a variant library
def dd_shrink(r: float, alpha: float, r0: float, rmin: float, L=None) -> float:
    v = ctypes.c_float(r)
    (L or lib()).oc_dd_shrink(ctypes.byref(v), alpha, r0, rmin)
    return v.value


def dd_grow(r: float, alpha: float, L=None) -> float:
    v = ctypes.c_float(r)
    (L or lib()).oc_dd_grow(ctypes.byref(v), alpha)
    return v.value


def dd_skips(radius: float, nd: float, L=None) -> bool:
    return bool((L or lib()).oc_dd_skips(radius, nd))


def balance_sides(mode: int, rnd: int, S: int, size_a: int, size_b: int, tree_ratio: float, vamp=None, L=None):
    """Tree of every explorer 0..S-1 in one round; ``vamp`` (mode-2 state per explorer) is
    updated in place."""
    vamp = np.zeros(S, np.int32) if vamp is None else vamp
    out = []
    for e in range(S):
        v = ctypes.c_int(int(vamp[e]))
        out.append((L or lib()).oc_balance_side(mode, rnd, e, S, size_a, size_b, tree_ratio, ctypes.byref(v)))
        vamp[e] = v.value
    return np.array(out, np.int32)


def steer(base, sample, nd: float, step: float, L=None) -> np.ndarray:
    b = np.ascontiguousarray(np.asarray(base, np.float32).reshape(-1))
    s = np.ascontiguousarray(np.asarray(sample, np.float32).reshape(-1))
    out = np.zeros_like(b)
    (L or lib()).oc_steer(_fp(b), _fp(s), b.shape[0], nd, step, _fp(out))
    return out


def connect_steps(frm, to, dist: float, step: float, L=None) -> tuple[int, np.ndarray]:
    f = np.ascontiguousarray(np.asarray(frm, np.float32).reshape(-1))
    t = np.ascontiguousarray(np.asarray(to, np.float32).reshape(-1))
    vec = np.zeros_like(f)
    n = (L or lib()).oc_connect_steps(_fp(f), _fp(t), f.shape[0], dist, step, _fp(vec))
    return int(n), vec


def plan(ctx: Context, start, goals, *, max_iterations=10_000, step_size=0.5, num_new_samples=128,
         granularity=16, max_nodes=16384, balance_mode=1, tree_ratio=0.5, dynamic_domain=True,
         dd_alpha=1e-4, dd_radius=4.0, dd_min_radius=1.0, min_vals=None, max_vals=None, seed=0, L=None) -> dict:
    """Run the lockstep planner; returns trees, parents, radii, sizes, iterations, conn, status."""
    dim = ctx.dim
    start = np.ascontiguousarray(np.asarray(start, np.float32).reshape(dim))
    goals = np.ascontiguousarray(np.asarray(goals, np.float32).reshape(-1, dim))
    lo = np.ascontiguousarray(np.full(dim, -np.pi, np.float32) if min_vals is None else np.asarray(min_vals, np.float32))
    hi = np.ascontiguousarray(np.full(dim, np.pi, np.float32) if max_vals is None else np.asarray(max_vals, np.float32))
    N = int(max_nodes)
    ta = np.zeros((dim, N), np.float32)
    tb = np.zeros((dim, N), np.float32)
    pa = np.full(N, -1, np.int32)
    pb = np.full(N, -1, np.int32)
    ra = np.zeros(N, np.float32)
    rb = np.zeros(N, np.float32)
    p = OcParams(int(max_iterations), float(step_size), int(num_new_samples), int(balance_mode), float(tree_ratio),
                 1 if dynamic_domain else 0, float(dd_alpha), float(dd_radius), float(dd_min_radius), dim, N,
                 int(granularity), ctypes.c_uint(seed).value)
    o = OcOutputs(_fp(ta), _fp(tb), _ip(pa), _ip(pb), _fp(ra), _fp(rb))
    rc = (L or lib()).oc_plan(ctypes.byref(ctx.c), ctypes.byref(p), _fp(start), _fp(goals), goals.shape[0],
                              _fp(lo), _fp(hi), ctypes.byref(o))
    if rc != 0:
        raise ValueError("oracle plan: invalid arguments")
    sa, sb = o.sizes[0], o.sizes[1]
    res = dict(tree_a=ta, tree_b=tb, parents_a=pa, parents_b=pb, radii_a=ra, radii_b=rb, sizes=(sa, sb),
               iterations=o.iterations, conn=tuple(o.conn), status=o.status, fkcc_evals=o.fkcc_evals)
    res["path"] = trace_path(res, L=L) if o.status == STATUS_SOLVED else None
    return res


def trace_path(res: dict, max_len: int | None = None, L=None) -> np.ndarray:
    ta, tb = res["tree_a"], res["tree_b"]
    dim, N = ta.shape
    max_len = max_len or (res["sizes"][0] + res["sizes"][1])
    path = np.zeros((max_len, dim), np.float32)
    conn = np.asarray(res["conn"], np.int32)
    n = (L or lib()).oc_trace_path(_fp(ta), _fp(tb), _ip(res["parents_a"]), _ip(res["parents_b"]), N, dim,
                                   _ip(conn), _fp(path), max_len)
    if n < 0:
        raise RuntimeError("path longer than max_len")
    return path[:n]


# --------------------------------------------------------------------------- FP64 restatement
def _skew(w):
    return np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])


def se3_exp64(twist, theta) -> np.ndarray:
    """4x4 exp of twist*theta, twist = (v, omega), by Rodrigues' formula."""
    v = np.asarray(twist[:3], np.float64) * theta
    w = np.asarray(twist[3:], np.float64) * theta
    th = float(np.linalg.norm(w))
    W = _skew(w)
    if th < 1e-12:
        R = np.eye(3) + W
        V = np.eye(3) + 0.5 * W
    else:
        A = math.sin(th) / th
        B = (1.0 - math.cos(th)) / th**2
        C = (th - math.sin(th)) / th**3
        R = np.eye(3) + A * W + B * (W @ W)
        V = np.eye(3) + B * W + C * (W @ W)
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = V @ v
    return T


def quat_mat64(q) -> np.ndarray:
    w, x, y, z = (float(a) for a in q)
    n = math.sqrt(w * w + x * x + y * y + z * z)
    w, x, y, z = w / n, x / n, y / n, z / n
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def fk64(cc: dict, q) -> np.ndarray:
    """FP64 FK: ``[n_joints, 4, 4]`` world transforms in original joint order."""
    return _fk64_sorted(cc, q)[np.asarray(cc["fk_topo_inv"])]


def _fk64_sorted(cc: dict, q) -> np.ndarray:
    """FP64 FK in the context's (topological) joint order."""
    q = np.asarray(q, np.float64)
    tw = np.asarray(cc["fk_twists"], np.float64)
    ptf = np.asarray(cc["fk_parent_tf"], np.float64)
    par = np.asarray(cc["fk_parent_idx"])
    act = np.asarray(cc["fk_act_idx"])
    mim = np.asarray(cc["fk_mimic_act_idx"])
    mul = np.asarray(cc["fk_mimic_mul"], np.float64)
    off = np.asarray(cc["fk_mimic_off"], np.float64)
    n = par.shape[0]
    T = np.zeros((n, 4, 4))
    for j in range(n):
        if mim[j] >= 0:
            th = q[mim[j]] * mul[j] + off[j]
        elif act[j] >= 0:
            th = q[act[j]]
        else:
            th = 0.0
        P = np.eye(4)
        P[:3, :3] = quat_mat64(ptf[j, :4])
        P[:3, 3] = ptf[j, 4:]
        base = T[par[j]] if par[j] >= 0 else np.eye(4)
        T[j] = base @ P @ se3_exp64(tw[j], th)
    return T


def sphere_positions64(cc: dict, q) -> np.ndarray:
    T = fk64(cc, q)
    links = np.asarray(cc["sphere_link_idx"])
    loc = np.asarray(cc["sphere_local"], np.float64)
    out = np.full((links.shape[0], 3), np.nan)
    n = T.shape[0]
    for s, l in enumerate(links):
        if 0 <= l < n:
            out[s] = T[l, :3, :3] @ loc[s] + T[l, :3, 3]
    return out


def pair_distances64(cc: dict, q) -> np.ndarray:
    """Every signed distance the binary predicate tests, in FP64."""
    P = sphere_positions64(cc, q)
    r = np.asarray(cc["sphere_radius"], np.float64)
    ok = ~np.isnan(P[:, 0])
    P, rr = P[ok], r[ok]
    d = []
    ws = np.asarray(cc.get("world_spheres", np.zeros((0, 4))), np.float64).reshape(-1, 4)
    if ws.shape[0]:
        d.append((np.linalg.norm(P[:, None, :] - ws[None, :, :3], axis=-1) - rr[:, None] - ws[None, :, 3]).ravel())
    wc = np.asarray(cc.get("world_capsules", np.zeros((0, 7))), np.float64).reshape(-1, 7)
    for c in wc:
        a, b = c[:3], c[3:6]
        ab = b - a
        t = np.clip(((P - a) @ ab) / max(ab @ ab, 1e-300), 0.0, 1.0)
        d.append(np.linalg.norm(P - (a + t[:, None] * ab), axis=-1) - rr - c[6])
    wb = np.asarray(cc.get("world_boxes", np.zeros((0, 15))), np.float64).reshape(-1, 15)
    for b in wb:
        R = np.stack([b[3:6], b[6:9], b[9:12]], axis=1)
        loc = (P - b[:3]) @ R
        qd = np.abs(loc) - b[12:15]
        outside = np.linalg.norm(np.maximum(qd, 0.0), axis=-1)
        inside = np.minimum(qd.max(axis=-1), 0.0)
        d.append(outside + inside - rr)
    wh = np.asarray(cc.get("world_halfspaces", np.zeros((0, 6))), np.float64).reshape(-1, 6)
    for h in wh:
        d.append((P - h[3:6]) @ h[:3] - rr)
    sp = np.asarray(cc.get("self_pairs", np.zeros((0, 2))), np.int64).reshape(-1, 2)
    if sp.shape[0]:
        Pfull = sphere_positions64(cc, q)
        n_rs = r.shape[0]
        good = (sp >= 0).all(1) & (sp < n_rs).all(1)
        sp = sp[good]
        good = ~np.isnan(Pfull[sp[:, 0], 0]) & ~np.isnan(Pfull[sp[:, 1], 0])
        sp = sp[good]
        d.append(np.linalg.norm(Pfull[sp[:, 0]] - Pfull[sp[:, 1]], axis=-1) - r[sp[:, 0]] - r[sp[:, 1]])
    return np.concatenate(d) if d else np.zeros(0)


def min_distance64(cc: dict, q) -> float:
    d = pair_distances64(cc, q)
    return float(d.min()) if d.size else float("inf")


# ---------------------------------------------------------------- sphere-penetration cost (FP64)
def _softplus64(z):
    return np.maximum(z, 0.0) + np.log1p(np.exp(-np.abs(z)))


def sphere_cost64(cc: dict, q, beta: float = 50.0) -> float:
    """sum over robot spheres i, world spheres w of softplus_beta(r_i + r_w - |c_i - o_w|), FP64."""
    P = sphere_positions64(cc, q)
    ok = ~np.isnan(P[:, 0])
    r = np.asarray(cc["sphere_radius"], np.float64)[ok]
    ws = np.asarray(cc.get("world_spheres", np.zeros((0, 4))), np.float64).reshape(-1, 4)
    if ws.shape[0] == 0 or not ok.any():
        return 0.0
    d = np.linalg.norm(P[ok][:, None, :] - ws[None, :, :3], axis=-1)
    x = r[:, None] + ws[None, :, 3] - d
    return float(_softplus64(beta * x).sum() / beta)


def sphere_cost_grad64(cc: dict, q, beta: float = 50.0) -> np.ndarray:
    """Analytic gradient of sphere_cost64 through the kinematic tree (adjoint form), FP64:
    sphere forces g_i, subtree sums F_j / M_j, then w_j . (M_j - t_j x F_j) + (R_j v_j) . F_j."""
    q = np.asarray(q, np.float64)
    Ts = _fk64_sorted(cc, q)
    n = Ts.shape[0]
    inv = np.asarray(cc["fk_topo_inv"])
    par = np.asarray(cc["fk_parent_idx"])
    links = np.asarray(cc["sphere_link_idx"])
    loc = np.asarray(cc["sphere_local"], np.float64)
    rad = np.asarray(cc["sphere_radius"], np.float64)
    ws = np.asarray(cc.get("world_spheres", np.zeros((0, 4))), np.float64).reshape(-1, 4)
    F = np.zeros((n, 3))
    M = np.zeros((n, 3))
    for s, l in enumerate(links):
        if not (0 <= l < n):
            continue
        j = inv[l]
        c = Ts[j, :3, :3] @ loc[s] + Ts[j, :3, 3]
        diff = c[None, :] - ws[:, :3]
        d = np.linalg.norm(diff, axis=1)
        z = beta * (rad[s] + ws[:, 3] - d)
        sig = 0.5 * (1.0 + np.tanh(0.5 * z))
        good = d > 0
        g = -(sig[good, None] * diff[good] / d[good, None]).sum(axis=0)
        F[j] += g
        M[j] += np.cross(c, g)
    for j in range(n - 1, -1, -1):
        if par[j] >= 0:
            F[par[j]] += F[j]
            M[par[j]] += M[j]
    tw = np.asarray(cc["fk_twists"], np.float64)
    gj = np.zeros(n)
    for j in range(n):
        R, t = Ts[j, :3, :3], Ts[j, :3, 3]
        w = R @ tw[j, 3:]
        v = R @ tw[j, :3]
        gj[j] = w @ (M[j] - np.cross(t, F[j])) + v @ F[j]
    act = np.asarray(cc["fk_act_idx"])
    mim = np.asarray(cc["fk_mimic_act_idx"])
    mul = np.asarray(cc["fk_mimic_mul"], np.float64)
    out = np.zeros(q.shape[0])
    for j in range(n):
        if mim[j] >= 0:
            out[mim[j]] += mul[j] * gj[j]
        elif act[j] >= 0:
            out[act[j]] += gj[j]
    return out
