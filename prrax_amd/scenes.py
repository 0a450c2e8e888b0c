"""Seeded synthetic robots, worlds and queries for the planner workloads.

Every array follows the reference ``collision_context`` wire format
(README.md:77-104, cuda-rrtc/jax/utils.py:119-252): FK arrays in joint (topological) order,
fine spheres sorted by FK frame, world primitives in the reference row layouts
(spheres ``[x,y,z,r]``, capsules ``[p1,p2,r]``, boxes ``[c,a1,a2,a3,hl]``, halfspaces
``[n,p]``).  All randomness is numpy PCG64 (``np.random.default_rng``) with fixed seeds, so a
workload is fully named by its builder and seed.

Workloads (sizes as quoted for the benchmark):

* ``c1`` -- 7-DOF serial arm, revolute axes z,y,z,y,z,y,z, link offsets U(0.10, 0.35) m along
  local z, limits +-2.8 rad, 48 spheres (7 per link, 6 on the last), 60 self pairs; world of
  40 spheres (r U(0.05, 0.15)) in a 1.2x1.2x1.0 m box in front of the base, plus a floor.
* ``c2`` -- same arm with 64 spheres / 200 self pairs; 200 spheres (r U(0.02, 0.12)), 16 oriented
  boxes, 8 capsules, floor; 4 goals.
* ``c3`` -- per query: 10-300 spheres, 0-24 boxes, floor, 1-8 goals, max_nodes 8192-65536 in
  proportion to obstacle count.
* ``c4`` -- 14-DOF dual arm (two ``c2`` arms 0.6 m apart) with a 2-finger gripper per arm whose
  finger joints mimic the wrist joint (mul 1, off 0): 18 joints, 128 spheres, ~2000 self pairs;
  world of 150 spheres, 12 boxes, floor.
* ``c5`` -- the ``c1`` arm against 100 world spheres (path refinement workload).

Start and goal configurations are rejection-sampled: ``sample_free`` takes a checker callback
``is_free(configs[n, dim]) -> bool[n]`` so that the same sampling runs against any checker.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

AXES_7 = ("z", "y", "z", "y", "z", "y", "z")
_AXIS = {"x": (1.0, 0.0, 0.0), "y": (0.0, 1.0, 0.0), "z": (0.0, 0.0, 1.0)}
LIMIT = 2.8
# Obstacle region in front of the base (x forward).
REGION_LO = np.array([0.3, -0.6, 0.0])
REGION_HI = np.array([1.5, 0.6, 1.0])


@dataclass
class Robot:
    """A robot in the reference FK/sphere wire format plus its joint limits."""

    ctx: dict
    lower: np.ndarray
    upper: np.ndarray

    @property
    def dim(self) -> int:
        return int(self.lower.shape[0])


@dataclass
class Scene:
    robot: Robot
    world: dict
    extra: dict = field(default_factory=dict)

    def context(self) -> dict:
        """Full ``collision_context`` (robot keys + world keys)."""
        out = dict(self.robot.ctx)
        out.update(self.world)
        return out


def _arm_chain(rng, n_spheres, axes=AXES_7):
    n = len(axes)
    lengths = rng.uniform(0.10, 0.35, size=n + 1)  # offsets of joints 0..n-1, then the tool
    per = [n_spheres // n] * n
    for k in range(n_spheres - sum(per)):
        per[k] += 1
    return lengths, per


def _arm_spheres(rng, per, lengths, r_lo, r_hi, link_base=0):
    links, local, radius = [], [], []
    for j, cnt in enumerate(per):
        zs = rng.uniform(0.0, lengths[j + 1], size=cnt)
        rs = rng.uniform(r_lo, r_hi, size=cnt)
        for z, r in zip(zs, rs):
            links.append(link_base + j)
            local.append((0.0, 0.0, z))
            radius.append(r)
    return links, local, radius


def _pick_pairs(rng, cand, n_pairs):
    cand = np.asarray(cand, dtype=np.int32).reshape(-1, 2)
    n_pairs = min(n_pairs, cand.shape[0])
    sel = np.sort(rng.choice(cand.shape[0], size=n_pairs, replace=False))
    return cand[sel]


def make_arm(seed: int = 0, n_spheres: int = 48, n_self_pairs: int = 60) -> Robot:
    """7-DOF serial arm (axes z,y,z,y,z,y,z), base at the origin."""
    rng = np.random.default_rng(seed)
    lengths, per = _arm_chain(rng, n_spheres)
    n = len(AXES_7)
    twists = np.zeros((n, 6), np.float32)
    parent_tf = np.zeros((n, 7), np.float32)
    for j, a in enumerate(AXES_7):
        twists[j, 3:] = _AXIS[a]
        parent_tf[j] = (1.0, 0.0, 0.0, 0.0, 0.0, 0.0, lengths[j])
    links, local, radius = _arm_spheres(rng, per, lengths, 0.04, 0.08)
    links = np.asarray(links, np.int32)
    cand = [(i, k) for i in range(len(links)) for k in range(i + 1, len(links))
            if abs(int(links[i]) - int(links[k])) >= 2]
    pairs = _pick_pairs(rng, cand, n_self_pairs)
    ctx = {
        "fk_twists": twists,
        "fk_parent_tf": parent_tf,
        "fk_parent_idx": np.arange(-1, n - 1, dtype=np.int32),
        "fk_act_idx": np.arange(n, dtype=np.int32),
        "fk_mimic_mul": np.ones(n, np.float32),
        "fk_mimic_off": np.zeros(n, np.float32),
        "fk_mimic_act_idx": np.full(n, -1, np.int32),
        "fk_topo_inv": np.arange(n, dtype=np.int32),
        "sphere_link_idx": links,
        "sphere_local": np.asarray(local, np.float32),
        "sphere_radius": np.asarray(radius, np.float32),
        "self_pairs": pairs.astype(np.int32),
    }
    return Robot(ctx, np.full(n, -LIMIT, np.float32), np.full(n, LIMIT, np.float32))


def make_dual_arm(seed: int = 0, n_self_pairs: int = 2000) -> Robot:
    """14-DOF dual arm: two 7-DOF chains 0.6 m apart, each with two mimic finger joints."""
    rng = np.random.default_rng(seed)
    lengths, per = _arm_chain(rng, 56)
    n_arm = 9  # 7 arm joints + 2 finger joints
    twists = np.zeros((2 * n_arm, 6), np.float32)
    parent_tf = np.zeros((2 * n_arm, 7), np.float32)
    parent = np.zeros(2 * n_arm, np.int32)
    act = np.full(2 * n_arm, -1, np.int32)
    mimic = np.full(2 * n_arm, -1, np.int32)
    links, local, radius, arm_of = [], [], [], []
    for arm, y in enumerate((0.3, -0.3)):
        b = arm * n_arm
        for j, a in enumerate(AXES_7):
            twists[b + j, 3:] = _AXIS[a]
            parent_tf[b + j] = (1.0, 0.0, 0.0, 0.0, 0.0, y if j == 0 else 0.0, lengths[j])
            parent[b + j] = -1 if j == 0 else b + j - 1
            act[b + j] = arm * 7 + j
        for f, fy in enumerate((0.03, -0.03)):
            jf = b + 7 + f
            twists[jf, 3:] = _AXIS["x"]
            parent_tf[jf] = (1.0, 0.0, 0.0, 0.0, 0.0, fy, lengths[7])
            parent[jf] = b + 6
            mimic[jf] = arm * 7 + 6
        l, lo, r = _arm_spheres(rng, per, lengths, 0.04, 0.08, link_base=b)
        links += l; local += lo; radius += r; arm_of += [arm] * len(l)
        for f in range(2):
            zs = np.linspace(0.01, 0.07, 4)
            for z in zs:
                links.append(b + 7 + f)
                local.append((0.0, 0.0, z))
                radius.append(0.012)
                arm_of.append(arm)
    links = np.asarray(links, np.int32)
    arm_of = np.asarray(arm_of)
    # link index within its arm: 0..6 arm links, 7 for both fingers
    rank = np.minimum(links % n_arm, 7)
    cand = []
    for i in range(len(links)):
        for k in range(i + 1, len(links)):
            if arm_of[i] != arm_of[k] or abs(int(rank[i]) - int(rank[k])) >= 2:
                cand.append((i, k))
    pairs = _pick_pairs(rng, cand, n_self_pairs)
    ctx = {
        "fk_twists": twists,
        "fk_parent_tf": parent_tf,
        "fk_parent_idx": parent,
        "fk_act_idx": act,
        "fk_mimic_mul": np.ones(2 * n_arm, np.float32),
        "fk_mimic_off": np.zeros(2 * n_arm, np.float32),
        "fk_mimic_act_idx": mimic,
        "fk_topo_inv": np.arange(2 * n_arm, dtype=np.int32),
        "sphere_link_idx": links,
        "sphere_local": np.asarray(local, np.float32),
        "sphere_radius": np.asarray(radius, np.float32),
        "self_pairs": pairs.astype(np.int32),
    }
    return Robot(ctx, np.full(14, -LIMIT, np.float32), np.full(14, LIMIT, np.float32))


def _random_quat(rng, n):
    q = rng.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def _quat_mat(q):
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def floor_for(robot: Robot) -> np.ndarray:
    """Floor halfspace ``[n, p]`` (n = +z) just below the lowest base-link sphere."""
    c = robot.ctx
    root = int(np.flatnonzero(c["fk_parent_idx"] < 0)[0])
    base_z = float(c["fk_parent_tf"][root, 6])
    sel = c["sphere_link_idx"] == root
    low = float(np.min(base_z + c["sphere_local"][sel, 2] - c["sphere_radius"][sel]))
    return np.array([[0.0, 0.0, 1.0, 0.0, 0.0, low - 0.01]], np.float32)


def root_spheres(robot: Robot) -> np.ndarray:
    """World ``[x, y, z, r]`` of the spheres on root frames whose twist is a rotation about the
    frame's own z axis through the origin: those spheres sit on the axis, so they never move."""
    c = robot.ctx
    out = []
    for j in np.flatnonzero(c["fk_parent_idx"] < 0):
        tw = c["fk_twists"][j]
        if np.any(tw[:3] != 0) or np.any(np.abs(tw[3:5]) > 0):
            continue
        R = _quat_mat(c["fk_parent_tf"][j, :4].astype(np.float64))
        t = c["fk_parent_tf"][j, 4:].astype(np.float64)
        for s in np.flatnonzero(c["sphere_link_idx"] == j):
            loc = c["sphere_local"][s].astype(np.float64)
            if abs(loc[0]) > 0 or abs(loc[1]) > 0:
                continue
            out.append(np.concatenate([R @ loc + t, [c["sphere_radius"][s]]]))
    return np.asarray(out, np.float64).reshape(-1, 4)


def _clear(kind: str, prim: np.ndarray, keep: np.ndarray, gap: float = 0.02) -> bool:
    """True when the primitive is at least ``gap`` away from every sphere in ``keep``."""
    if keep.shape[0] == 0:
        return True
    c, r = keep[:, :3], keep[:, 3]
    if kind == "sphere":
        d = np.linalg.norm(c - prim[:3], axis=1) - prim[3]
    elif kind == "capsule":
        a, b = prim[:3], prim[3:6]
        ab = b - a
        t = np.clip((c - a) @ ab / max(ab @ ab, 1e-300), 0.0, 1.0)
        d = np.linalg.norm(c - (a + t[:, None] * ab), axis=1) - prim[6]
    else:
        R = np.stack([prim[3:6], prim[6:9], prim[9:12]], axis=1)
        q = np.abs((c - prim[:3]) @ R) - prim[12:15]
        d = np.linalg.norm(np.maximum(q, 0.0), axis=1) + np.minimum(q.max(axis=1), 0.0)
    return bool(np.all(d - r >= gap))


def make_world(rng, n_spheres, r_lo, r_hi, n_boxes=0, n_capsules=0, floor=None, keep_clear=None) -> dict:
    """World primitives in the obstacle region; an obstacle that would touch a sphere of
    ``keep_clear`` (the robot's immovable base) has its centre redrawn."""
    keep = np.zeros((0, 4)) if keep_clear is None else np.asarray(keep_clear, np.float64)
    spheres = np.zeros((n_spheres, 4), np.float32)
    spheres[:, :3] = rng.uniform(REGION_LO, REGION_HI, size=(n_spheres, 3))
    spheres[:, 3] = rng.uniform(r_lo, r_hi, size=n_spheres)
    for k in range(n_spheres):
        while not _clear("sphere", spheres[k].astype(np.float64), keep):
            spheres[k, :3] = rng.uniform(REGION_LO, REGION_HI)
    boxes = np.zeros((n_boxes, 15), np.float32)
    if n_boxes:
        centers = rng.uniform(REGION_LO, REGION_HI, size=(n_boxes, 3))
        hl = rng.uniform(0.05, 0.3, size=(n_boxes, 3))
        quats = _random_quat(rng, n_boxes)
        for k in range(n_boxes):
            R = _quat_mat(quats[k])
            boxes[k] = np.concatenate([centers[k], R[:, 0], R[:, 1], R[:, 2], hl[k]])
            while not _clear("box", boxes[k].astype(np.float64), keep):
                boxes[k, :3] = rng.uniform(REGION_LO, REGION_HI)
    caps = np.zeros((n_capsules, 7), np.float32)
    if n_capsules:
        centers = rng.uniform(REGION_LO, REGION_HI, size=(n_capsules, 3))
        length = rng.uniform(0.2, 0.6, size=n_capsules)
        rad = rng.uniform(0.03, 0.08, size=n_capsules)
        d = rng.normal(size=(n_capsules, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        caps[:, 0:3] = centers - d * (length[:, None] / 2)
        caps[:, 3:6] = centers + d * (length[:, None] / 2)
        caps[:, 6] = rad
        for k in range(n_capsules):
            while not _clear("capsule", caps[k].astype(np.float64), keep):
                c = rng.uniform(REGION_LO, REGION_HI)
                caps[k, 0:3] = c - d[k] * length[k] / 2
                caps[k, 3:6] = c + d[k] * length[k] / 2
    return {
        "world_spheres": spheres,
        "world_capsules": caps,
        "world_boxes": boxes,
        "world_halfspaces": floor if floor is not None else np.zeros((0, 6), np.float32),
    }


def scene_c1(robot_seed: int = 0, world_seed: int = 1) -> Scene:
    robot = make_arm(robot_seed, 48, 60)
    world = make_world(np.random.default_rng(world_seed), 40, 0.05, 0.15, floor=floor_for(robot), keep_clear=root_spheres(robot))
    return Scene(robot, world)


def scene_c2(robot_seed: int = 0, world_seed: int = 1) -> Scene:
    robot = make_arm(robot_seed, 64, 200)
    world = make_world(np.random.default_rng(world_seed), 200, 0.02, 0.12, n_boxes=16, n_capsules=8,
                       floor=floor_for(robot), keep_clear=root_spheres(robot))
    return Scene(robot, world)


def scene_c4(robot_seed: int = 0, world_seed: int = 1) -> Scene:
    robot = make_dual_arm(robot_seed)
    world = make_world(np.random.default_rng(world_seed), 150, 0.02, 0.12, n_boxes=12, floor=floor_for(robot), keep_clear=root_spheres(robot))
    return Scene(robot, world)


def scene_c5(robot_seed: int = 0, world_seed: int = 1) -> Scene:
    robot = make_arm(robot_seed, 48, 60)
    world = make_world(np.random.default_rng(world_seed), 100, 0.05, 0.15, floor=floor_for(robot), keep_clear=root_spheres(robot))
    return Scene(robot, world)


C3_SPHERES = (10, 300)
C3_BOXES = (0, 24)
C3_GOALS = (1, 8)
C3_NODES = (8192, 65536)


def c3_query_world(robot: Robot, query: int, seed_base: int = 3000) -> tuple[dict, int, int]:
    """World, goal count and max_nodes of query ``query`` of the 1024-query batch."""
    rng = np.random.default_rng(seed_base + query)
    ns = int(rng.integers(C3_SPHERES[0], C3_SPHERES[1] + 1))
    nb = int(rng.integers(C3_BOXES[0], C3_BOXES[1] + 1))
    ng = int(rng.integers(C3_GOALS[0], C3_GOALS[1] + 1))
    world = make_world(rng, ns, 0.02, 0.12, n_boxes=nb, floor=floor_for(robot), keep_clear=root_spheres(robot))
    n_min = C3_SPHERES[0] + C3_BOXES[0]
    n_max = C3_SPHERES[1] + C3_BOXES[1]
    frac = (ns + nb - n_min) / (n_max - n_min)
    nodes = C3_NODES[0] + (C3_NODES[1] - C3_NODES[0]) * frac
    nodes = int(round(nodes / 64.0)) * 64
    return world, ng, max(C3_NODES[0], min(C3_NODES[1], nodes))


def sample_free(rng, robot: Robot, n: int, is_free, batch: int = 256, max_tries: int = 1_000_000):
    """``n`` configurations drawn from U(limits) and kept when ``is_free`` accepts them."""
    out = []
    tried = 0
    while len(out) < n:
        if tried >= max_tries:
            raise RuntimeError(f"sample_free: only {len(out)}/{n} free configs after {tried} draws")
        q = rng.uniform(robot.lower, robot.upper, size=(batch, robot.dim)).astype(np.float32)
        tried += batch
        ok = np.asarray(is_free(q), dtype=bool)
        for k in np.flatnonzero(ok):
            out.append(q[k])
            if len(out) == n:
                break
    return np.stack(out, axis=0)


def query_seed(config: int, query: int) -> int:
    """Seed of start/goal sampling for query ``query`` of workload ``config``."""
    return 1000 * config + query


def scene_cloud(n_points: int = 100_000, robot_seed: int = 0, world_seed: int = 1, r_env: float = 0.02) -> Scene:
    """Point-cloud scene: the C1 arm, the C1 world's 40 spheres sampled as ``n_points`` surface
    points (uniform over the total area; a dense capture of the same objects), and the floor
    halfspace as the analytic world.  ``extra`` holds ``points`` [n, 3] and ``r_env``."""
    sc = scene_c1(robot_seed, world_seed)
    rng = np.random.default_rng(world_seed + 7000)
    ws = sc.world["world_spheres"].astype(np.float64)
    area = ws[:, 3] ** 2
    k = rng.choice(ws.shape[0], size=n_points, p=area / area.sum())
    d = rng.normal(size=(n_points, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pts = ws[k, :3] + d * ws[k, 3:4]
    world = {"world_halfspaces": sc.world["world_halfspaces"]}
    return Scene(sc.robot, world, extra={"points": pts.astype(np.float32), "r_env": float(r_env)})
