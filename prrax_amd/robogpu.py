"""Point-cloud edge validation and the host-driven planner that uses it.

The reference's ``prrtc_plan_robogpu`` (cuda-rrtc/jax/prrtc_robogpu.py:103-298) runs batched
RRT-Connect on the host and validates every round's candidate edges in one call of pyroffi's
OptiX point-cloud checker (``RoboGPUCollisionChecker.check_edges_collision_free``,
prrtc_robogpu.py:186-188): the analytic world geometry first, then the environment cloud, each
point a sphere of radius ``r_env``.  Here the checker is the library's ``prrtc_check_edges``: the
planner's own culled world verdict, then a uniform grid over the cloud (``prrtc_cloud_build``)
that each robot sphere queries for the few cells its reach overlaps.

``PointCloudChecker`` stands where the reference's ``RoboGPUCollisionChecker`` stood and
``prrtc_plan_robogpu`` keeps the reference's signature and loop (prrtc_robogpu.py:183-264):
``robot`` provides the joint limits (``lower``/``upper``, or ``joints.lower_limits`` /
``upper_limits``) and the checker holds the robot's FK and sphere model.  As in the reference,
the cloud's grid is built once per cloud object and reused across calls.
"""

from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np
import torch

from . import _lib
from .context import as_numpy, check_context, upload_robot, upload_world
from .prrtc import PRRTCResult, _device, _resolve_collision_mode, _stream_ptr

WORLD_KEYS = ("world_spheres", "world_capsules", "world_boxes", "world_halfspaces")
MAX_CELLS = 1 << 22


class PointCloud:
    """An environment cloud ``[n, 3]`` in the library's device grid.

    The cell edge defaults to half the largest robot-sphere reach (radius + ``r_env``), enlarged
    until the grid has at most 2^22 cells: a sphere's query then visits about five rows of cells
    per axis, each clipped to the sphere's chord, so the points it tests lie close to its reach."""

    def __init__(self, points, r_env: float, max_robot_radius: float, lo=None, hi=None, cell: Optional[float] = None):
        dev = _device()
        pts = points if isinstance(points, torch.Tensor) else torch.from_numpy(as_numpy(points, np.float32))
        self.points_in = pts.to(dev, torch.float32).reshape(-1, 3).contiguous()
        n = int(self.points_in.shape[0])
        if n:
            mn, mx = torch.aminmax(self.points_in, dim=0)
            lo = mn.cpu().numpy() if lo is None else np.asarray(lo, np.float32)
            hi = mx.cpu().numpy() if hi is None else np.asarray(hi, np.float32)
        else:
            lo = np.zeros(3, np.float32) if lo is None else np.asarray(lo, np.float32)
            hi = lo.copy() if hi is None else np.asarray(hi, np.float32)
        self.lo = np.ascontiguousarray(lo, np.float32)
        self.hi = np.ascontiguousarray(np.maximum(hi, lo), np.float32)
        if cell is None:
            cell = max(0.5 * (float(max_robot_radius) + float(r_env)), 1e-3)
            ext = np.maximum(self.hi - self.lo, 0.0).astype(np.float64)
            while np.prod(np.maximum(np.ceil(ext / cell), 1.0)) > MAX_CELLS:
                cell *= 1.25
        self.cell = float(cell)
        self.r_env = float(r_env)
        f3 = ctypes.POINTER(ctypes.c_float)
        lo_p, hi_p = self.lo.ctypes.data_as(f3), self.hi.ctypes.data_as(f3)
        nbytes = _lib.lib().prrtc_cloud_bytes(n, lo_p, hi_p, self.cell)
        if nbytes == 0:
            msg = _lib.lib().prrtc_last_error()
            raise ValueError(f"point cloud grid rejected: {msg.decode() if msg else 'bad bounds or cell'}")
        self.memory = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
        self.struct = _lib.Cloud()
        rc = _lib.lib().prrtc_cloud_build(_stream_ptr(), self.points_in.data_ptr() if n else None, n, lo_p, hi_p,
                                          self.cell, self.r_env, self.memory.data_ptr(), int(nbytes),
                                          ctypes.byref(self.struct))
        _lib.check("prrtc_cloud_build", rc)
        self.n_points = n


class PointCloudChecker:
    """Edge validation against world geometry, the robot's self pairs and a point cloud
    (the role of pyroffi's ``RoboGPUCollisionChecker``).

    ``collision_context`` gives the robot keys (FK, spheres, ``self_pairs``); world keys in it
    are ignored: the world comes with each call, as in the reference."""

    def __init__(self, collision_context: dict, dim: int, collision_checker: str = "binary",
                 collision_margin: float = 0.02, cloud_cell: Optional[float] = None):
        self.dim = int(dim)
        self.cloud_cell = cloud_cell  # grid cell edge; None = PointCloud's default
        self.mode = _resolve_collision_mode(collision_checker)
        self.margin = float(collision_margin)
        robot_cc = {k: v for k, v in collision_context.items() if k not in WORLD_KEYS}
        self.robot_cc = robot_cc
        self.arrays = check_context(robot_cc, self.dim)
        self.device = _device()
        self.robot = upload_robot(self.arrays, self.device)
        rad = self.arrays["sphere_radius"]
        self.max_radius = float(np.max(rad)) if rad.size else 0.0
        self._world_key = None
        self._world = None
        self._cloud_key = None
        self._cloud = None
        self.launch = _lib.Launch(0, 0, -1, 0, 0, 0)

    def _world_for(self, world_geom):
        key = None if world_geom is None else id(world_geom)
        if self._world is None or key != self._world_key:
            cc = dict(self.robot_cc)
            for k in WORLD_KEYS:
                if world_geom is not None and world_geom.get(k) is not None:
                    cc[k] = world_geom[k]
            self._world = upload_world(check_context(cc, self.dim), self.device)
            self._world_key = key
            self._world_ref = world_geom  # keeps the keyed object alive, so its id stays unique
        return self._world

    def cloud_for(self, point_cloud, r_env: float) -> Optional[PointCloud]:
        """The grid of ``point_cloud``, built on first use and cached on the object's identity."""
        if point_cloud is None:
            return None
        shape = tuple(point_cloud.shape)
        ptr = point_cloud.data_ptr() if isinstance(point_cloud, torch.Tensor) else None
        key = (id(point_cloud), ptr, shape, float(r_env))
        if self._cloud is None or key != self._cloud_key:
            self._cloud = PointCloud(point_cloud, r_env, self.max_radius, cell=self.cloud_cell)
            self._cloud_key = key
            self._cloud_ref = point_cloud  # keeps the keyed object alive, so its id stays unique
        return self._cloud

    def edges_free(self, edges, world_geom=None, point_cloud=None, r_env: float = 0.02,
                   flops: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Device uint8 ``[S]``: 1 iff all ``G`` waypoints of edge s (``edges [S, G, dim]``) are free."""
        if isinstance(edges, torch.Tensor):
            wp = edges.to(self.device, torch.float32).contiguous()
        else:
            wp = torch.from_numpy(as_numpy(edges, np.float32)).to(self.device)
        if wp.dim() != 3 or wp.shape[2] != self.dim:
            raise ValueError(f"edges must have shape (S, G, {self.dim}), got {tuple(wp.shape)}")
        S, G = int(wp.shape[0]), int(wp.shape[1])
        out = torch.empty(S, dtype=torch.uint8, device=self.device)
        if S == 0:
            return out
        world = self._world_for(world_geom)
        cloud = self.cloud_for(point_cloud, r_env)
        rc = _lib.lib().prrtc_check_edges(
            _stream_ptr(), wp.data_ptr(), S, G, self.dim, self.robot.struct, world.struct, self.mode, self.margin,
            None if cloud is None or cloud.n_points == 0 else ctypes.byref(cloud.struct), out.data_ptr(), self.launch,
            None if flops is None else flops.data_ptr())
        _lib.check("prrtc_check_edges", rc)
        return out

    def check_edges_collision_free(self, robot, edges, world_geom=None, point_cloud=None, r_env: float = 0.02):
        """The reference's call (prrtc_robogpu.py:186-188): ``[S]`` int32 on the host, 1 = free.
        ``robot`` is accepted for the signature; the checker's own robot model is used."""
        return self.edges_free(edges, world_geom, point_cloud, r_env).cpu().numpy().astype(np.int32)


def _limits_of(robot, dim: int, vals, which: str) -> np.ndarray:
    if vals is not None:
        return np.asarray(vals, np.float32).reshape(dim)
    lim = getattr(robot, "lower" if which == "lo" else "upper", None)
    if lim is None:
        joints = getattr(robot, "joints", None)
        lim = getattr(joints, "lower_limits" if which == "lo" else "upper_limits", None)
    if lim is not None:
        return np.asarray(lim, np.float32).reshape(dim)
    return np.full(dim, -np.pi if which == "lo" else np.pi, np.float32)


class _HostTree:
    """Preallocated host tree (rows of configurations, parents); root at index 0, its own parent."""

    def __init__(self, root: np.ndarray, max_nodes: int, dim: int):
        self.configs = np.zeros((max_nodes, dim), np.float32)
        self.parents = np.full(max_nodes, -1, np.int32)
        self.configs[0] = root
        self.parents[0] = 0
        self.size = 1

    def room(self) -> int:
        return self.configs.shape[0] - self.size

    def add_many(self, configs: np.ndarray, parents: np.ndarray) -> np.ndarray:
        """Appends rows in order until the tree is full; returns the indices of those that fit
        (the reference adds one node at a time and stops at the first that does not fit)."""
        n = min(configs.shape[0], self.configs.shape[0] - self.size)
        i0 = self.size
        self.configs[i0:i0 + n] = configs[:n]
        self.parents[i0:i0 + n] = parents[:n]
        self.size += n
        return np.arange(i0, i0 + n, dtype=np.int64)

    def nearest(self, queries: np.ndarray, chunk: int = 8192) -> np.ndarray:
        """Lowest-index nearest node of each query.  Per (node, query) the squared distance is the
        reference's ``einsum`` over dimensions; node chunks bound the temporary, and a strict
        comparison keeps the earlier chunk on ties, as one ``argmin`` would."""
        best = np.full(queries.shape[0], np.inf, np.float32)
        arg = np.zeros(queries.shape[0], np.int64)
        for k0 in range(0, self.size, chunk):
            v = self.configs[k0:min(self.size, k0 + chunk)]
            diff = v[:, None, :] - queries[None, :, :]
            d2 = np.einsum("ksd,ksd->ks", diff, diff)
            m = d2.min(axis=0)
            better = m < best
            arg[better] = d2.argmin(axis=0)[better] + k0
            best[better] = m[better]
        return arg

    def path_to(self, node: int) -> list:
        chain = []
        cur = int(node)
        for _ in range(self.size):
            chain.append(self.configs[cur])
            par = int(self.parents[cur])
            if par == cur:
                break
            cur = par
        chain.reverse()
        return chain


def _steer(q_from: np.ndarray, q_to: np.ndarray, step_size: float) -> np.ndarray:
    delta = q_to - q_from
    dist = np.linalg.norm(delta, axis=1, keepdims=True)
    return q_from + delta * np.minimum(1.0, step_size / np.maximum(dist, 1e-8))


def prrtc_plan_robogpu(start_config, goal_configs, robot, robogpu_checker: PointCloudChecker, world_geom=None,
                       point_cloud=None, r_env: float = 0.02, max_iterations: int = 5_000, step_size: float = 0.5,
                       num_new_samples: int = 128, granularity: int = 16, max_nodes: int = 1_000_000,
                       connect_max_steps: int = 32, min_vals=None, max_vals=None, seed: int = 0) -> PRRTCResult:
    """Batched RRT-Connect driven from the host with point-cloud edge validation on the GPU
    (prrtc_robogpu.py:103-298: the same arguments, loop and result).

    Each iteration samples ``num_new_samples`` configurations, extends the active tree toward
    them by at most ``step_size``, validates all extension edges (``granularity`` waypoints at
    ``linspace(0, 1, G)``, end points included) in one device call, inserts the free ones, then
    greedily connects the other tree toward each new node for up to ``connect_max_steps`` steps;
    a step landing within ``step_size * 1e-3 + 1e-6`` of its target connects the trees.  The
    trees swap roles every iteration; the goal tree is rooted at the first goal."""
    start = np.asarray(as_numpy(start_config, np.float32)).reshape(-1)
    dim = start.shape[0]
    goal = np.asarray(as_numpy(goal_configs, np.float32)).reshape(-1, dim)[0]
    lo = _limits_of(robot, dim, min_vals, "lo")
    hi = _limits_of(robot, dim, max_vals, "hi")
    rng = np.random.default_rng(seed)
    tree_a = _HostTree(start, max_nodes, dim)
    tree_b = _HostTree(goal, max_nodes, dim)
    ts = np.linspace(0.0, 1.0, granularity, dtype=np.float32)
    t_kernel = [0.0]

    def edges_free(qa: np.ndarray, qb: np.ndarray) -> np.ndarray:
        if qa.shape[0] == 0:
            return np.zeros(0, bool)
        wp = (qa[:, None, :] * (1.0 - ts)[None, :, None] + qb[:, None, :] * ts[None, :, None]).astype(np.float32)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        free = robogpu_checker.edges_free(wp, world_geom, point_cloud, r_env)
        ev1.record()
        out = free.cpu().numpy() == 1
        t_kernel[0] += ev0.elapsed_time(ev1)
        return out

    solved = False
    connection = None
    iterations = 0
    active, other = tree_a, tree_b
    for iterations in range(1, max_iterations + 1):
        samples = rng.uniform(lo, hi, size=(num_new_samples, dim)).astype(np.float32)
        near_idx = active.nearest(samples)
        q_near = active.configs[:active.size][near_idx]
        q_new = _steer(q_near, samples, step_size)
        fs = np.nonzero(edges_free(q_near, q_new))[0]
        active_ids = active.add_many(q_new[fs], near_idx[fs])
        if active_ids.size == 0:
            active, other = other, active
            continue
        targets = q_new[fs[:active_ids.size]]
        other_idx = other.nearest(targets)
        n = targets.shape[0]
        reach = step_size * 1e-3 + 1e-6
        # The walks' positions on the host: a step that is taken starts where the previous one
        # ended, so they do not depend on the verdicts.  A walk has at most connect_max_steps
        # steps and ends at the first step within reach of its target.  Its steps are checked
        # in blocks of 2, 4, 8, ... steps of the walks still live, one device call per block
        # (positions computed as far as the block needs), and replayed in the reference's order.
        pos = [other.configs[:other.size][other_idx]]
        reach_step = np.zeros(n, np.int64)  # step that comes within reach (0: none so far)

        def walk_to(upto: int) -> None:
            while len(pos) <= upto and not (reach_step > 0).all():
                j = len(pos)
                pos.append(_steer(pos[-1], targets, step_size))
                diff = pos[j] - targets
                now = (reach_step == 0) & (np.sqrt(np.einsum("kd,kd->k", diff, diff)) <= reach)
                reach_step[now] = j

        free_grid = np.zeros((n, connect_max_steps), bool)
        live = np.ones(n, bool)
        checked, block = 0, 2
        for j in range(1, connect_max_steps + 1):
            if not live.any():
                break
            if j > checked:
                walk_to(min(connect_max_steps, checked + block))
                checked = len(pos) - 1
                block *= 2
                if j > checked:
                    break  # every walk reached its target at an earlier step
                lv = np.nonzero(live)[0]
                last = np.where(reach_step[lv] > 0, np.minimum(reach_step[lv], checked), checked)
                tt, jj = np.nonzero(np.arange(j, checked + 1)[None, :] <= last[:, None])
                tt, jj = lv[tt], jj + (j - 1)
                P0 = np.stack(pos[j - 1:checked], axis=1)  # [n, checked - j + 1, dim]
                P1 = np.stack(pos[j:checked + 1], axis=1)
                free_grid[tt, jj] = edges_free(P0[tt, jj - (j - 1)], P1[tt, jj - (j - 1)])
            li = np.nonzero(live)[0]
            free_c = free_grid[li, j - 1]
            # The reference walks li in order: a blocked step ends its target, a free one is
            # added (until the tree is full) and the first added step within reach of its
            # target connects the trees, leaving the rest of li unvisited.
            live[li[~free_c]] = False
            ok = np.nonzero(free_c)[0]
            fit = ok[:other.room()]
            hit = np.nonzero(reach_step[li[fit]] == j)[0]
            if hit.size:
                fit = fit[:hit[0] + 1]
            ids = other.add_many(pos[j][li[fit]], other_idx[li[fit]])
            other_idx[li[fit]] = ids
            if not hit.size:
                live[li[ok[fit.size:]]] = False
            else:
                s = int(li[fit[-1]])
                live[s] = False
                a_node, b_node = int(active_ids[s]), int(ids[-1])
                connection = (a_node, b_node) if active is tree_a else (b_node, a_node)
                solved = True
                break
        if solved:
            break
        active, other = other, active

    na, nb = tree_a.size, tree_b.size
    if solved:
        a_node, b_node = connection
        pa = tree_a.path_to(a_node)
        pb = tree_b.path_to(b_node)
        pb.reverse()
        path = np.stack(pa + pb, axis=0)
        cost = float(np.sum(np.linalg.norm(np.diff(path, axis=0), axis=1)))
    else:
        path, cost = None, float("inf")
    return PRRTCResult(solved=solved, path=path, tree_a_size=na, tree_b_size=nb, iterations=iterations, cost=cost,
                       kernel_time_ms=t_kernel[0], tree_a_configs=tree_a.configs[:na].T.copy(),
                       tree_b_configs=tree_b.configs[:nb].T.copy(), tree_a_parents=tree_a.parents[:na].copy(),
                       tree_b_parents=tree_b.parents[:nb].copy())
