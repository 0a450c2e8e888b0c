"""GPU: the point-cloud grid, edge validation against it, and the host-driven point-cloud planner.

* The grid built on the device equals a NumPy build with the same float32 cell arithmetic
  (cell starts exactly; each cell's points as a set).
* ``prrtc_check_edges`` gives, per edge, exactly the oracle's verdict: no waypoint in collision
  with the world (the planner's predicate, bit-exact) and none within r + r_env of a cloud point
  (brute force over every point), on edges near contact, in both modes, with and without a
  cloud, for several granularities and grid cell sizes.
* ``prrtc_plan_robogpu`` equals the reference loop restated on the CPU (``oracle.host_rrtc``)
  with the oracle as its checker: the same trees, iterations and path, seed for seed; the path
  is free under an FP64 check with the cloud's points as spheres.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import host_rrtc
from oracle import oracle as O
from prrax_amd import scenes, validate
from prrax_amd.robogpu import PointCloud, PointCloudChecker, prrtc_plan_robogpu

from test_simt_bounds import edge_configs

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _grid_numpy(pts, lo, cell, shape):
    inv = np.float32(1.0) / np.float32(cell)
    idx = []
    for a in range(3):
        f = np.floor((pts[:, a] - np.float32(lo[a])) * inv)
        idx.append(np.clip(f, 0, shape[a] - 1).astype(np.int64))
    return (idx[2] * shape[1] + idx[1]) * shape[0] + idx[0]


@pytest.mark.parametrize("n,cell", [(20_000, None), (5_000, 0.02), (1, None), (0, None)])
def test_grid_build_matches_numpy(n, cell):
    say(f"grid build n={n} cell={cell}")
    sc = scenes.scene_cloud(max(n, 1))
    pts = sc.extra["points"][:n]
    pc = PointCloud(pts, 0.02, 0.08, cell=cell)
    s = pc.struct
    shape = (s.nx, s.ny, s.nz)
    cells = s.nx * s.ny * s.nz
    mem = pc.memory.cpu().numpy()
    start = np.frombuffer(mem.tobytes(), np.int32, cells + 1, s.cell_start - pc.memory.data_ptr())
    if n == 0:
        assert start[-1] == 0
        return
    cid = _grid_numpy(pts, np.array(s.lo[:]), s.cell, shape)
    want = np.concatenate([[0], np.cumsum(np.bincount(cid, minlength=cells))])
    np.testing.assert_array_equal(start, want)
    sorted_pts = np.frombuffer(mem.tobytes(), np.float32, 4 * n, s.points - pc.memory.data_ptr()).reshape(n, 4)
    order = np.argsort(cid, kind="stable")
    for c in np.unique(cid)[:200]:
        a = {tuple(r) for r in sorted_pts[start[c]:start[c + 1], :3]}
        b = {tuple(r) for r in pts[order][want[c]:want[c + 1]]}
        assert a == b


def _oracle_edges_free(cc, dim, wp, pts, r_env, mode=0, margin=0.0):
    S, G = wp.shape[:2]
    q = wp.reshape(-1, dim)
    hit = O.in_collision(O.Context(cc, dim, collision_mode=mode, collision_margin=margin), q)
    if pts is not None and len(pts):
        hit |= O.cloud_hits(O.Context(cc, dim), q, pts, r_env)
    return ~hit.reshape(S, G).any(axis=1)


@pytest.mark.parametrize("G,mode,cloud,cell", [(16, 0, True, None), (16, 1, True, None), (5, 0, True, 0.03),
                                                (1, 0, True, None), (16, 0, False, None)])
def test_check_edges_matches_oracle(G, mode, cloud, cell):
    say(f"check edges G={G} mode={mode} cloud={cloud} cell={cell}")
    sc = scenes.scene_cloud(20_000)
    pts, r_env = sc.extra["points"], sc.extra["r_env"]
    cc = sc.context()
    rng = np.random.default_rng(20 + G)
    a = edge_configs(rng, sc.robot.lower, sc.robot.upper, 32, 0.3)[::16][:, None, :]  # 32 edge starts
    b = a + rng.normal(scale=0.15, size=a.shape).astype(np.float32)
    t = np.linspace(0, 1, G, dtype=np.float32)[None, :, None]
    wp = np.concatenate([(a * (1 - t) + b * t).astype(np.float32),
                         rng.uniform(sc.robot.lower, sc.robot.upper, size=(64, G, 7)).astype(np.float32)])
    chk = PointCloudChecker(cc, 7, collision_checker="differentiable" if mode else "binary", collision_margin=0.01,
                            cloud_cell=cell)
    pcl = torch.from_numpy(pts).cuda() if cloud else None
    got = chk.check_edges_collision_free(None, wp, world_geom=sc.world, point_cloud=pcl, r_env=r_env) == 1
    want = _oracle_edges_free(cc, 7, wp, pts if cloud else None, r_env, mode, 0.01)
    np.testing.assert_array_equal(got, want)
    assert 0.05 < want.mean() < 0.95


class _OracleCloudChecker:
    """The oracle as the host loop's checker: world predicate or cloud contact."""

    def __init__(self, cc, dim, pts, r_env):
        self.dim, self.ctx, self.pts, self.r_env = dim, O.Context(cc, dim), pts, r_env

    def free(self, qs):
        qs = np.asarray(qs, np.float32).reshape(-1, self.dim)
        return ~(O.in_collision(self.ctx, qs) | O.cloud_hits(self.ctx, qs, self.pts, self.r_env))


def test_plan_robogpu_equals_the_reference_loop():
    sc = scenes.scene_cloud(3000)
    pts, r_env = sc.extra["points"], sc.extra["r_env"]
    cc = sc.context()
    robot = sc.robot
    chk = PointCloudChecker(cc, 7)
    pcl = torch.from_numpy(pts).cuda()
    ora = _OracleCloudChecker(cc, 7, pts, r_env)
    free_fn = lambda q: ora.free(q)  # noqa: E731
    for qi in range(3):
        sg = scenes.sample_free(np.random.default_rng(scenes.query_seed(6, qi)), robot, 2, free_fn)
        say(f"plan robogpu query {qi}")
        kw = dict(max_iterations=300, step_size=0.5, num_new_samples=32, granularity=16, max_nodes=20_000,
                  connect_max_steps=32, seed=qi)
        res = prrtc_plan_robogpu(sg[0], sg[1:], robot, chk, world_geom=sc.world, point_cloud=pcl, r_env=r_env, **kw)
        ref = host_rrtc.plan(ora, sg[0], sg[1:], min_vals=robot.lower, max_vals=robot.upper, **kw)
        assert res.solved == ref["solved"] and res.iterations == ref["iterations"]
        assert (res.tree_a_size, res.tree_b_size) == ref["sizes"]
        if res.solved:
            np.testing.assert_array_equal(res.path, ref["path"])
            world = dict(cc, world_spheres=np.concatenate([pts, np.full((len(pts), 1), r_env, np.float32)], 1))
            assert validate.validate_paths(world, [res.path])[0]
