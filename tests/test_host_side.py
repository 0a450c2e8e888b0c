"""Host-side pieces on the CPU: the FP64 path validator, the NumPy restatement of the reference's
host planner loop, the C-ABI struct layouts and the context checks of the coarse keys."""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import host_rrtc
from oracle import oracle as O
from prrax_amd import _lib, scenes, validate
from prrax_amd.context import check_context

from helpers import oracle_free_fn, query
from test_oracle_fk_collision import random_tree_robot

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.parametrize("name", ["c1", "c2", "c4", "tree"])
def test_validator_matches_oracle_fp64(name):
    """Two independent FP64 restatements (the validator, vectorised over configurations, and the
    oracle's per-configuration one) agree to rounding, also with the screening cap."""
    rng = np.random.default_rng(4)
    if name == "tree":
        cc, dim = random_tree_robot(3)
        cc["world_spheres"] = rng.uniform(-0.5, 0.5, (30, 4)).astype(np.float32)
        cc["world_boxes"] = np.array([[0.1, 0, 0.2, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0.1, 0.2, 0.05]], np.float32)
        lo, hi = np.full(dim, -2.8), np.full(dim, 2.8)
    else:
        sc = {"c1": scenes.scene_c1, "c2": scenes.scene_c2, "c4": scenes.scene_c4}[name]()
        cc, dim, lo, hi = sc.context(), sc.robot.dim, sc.robot.lower, sc.robot.upper
    q = rng.uniform(lo, hi, size=(150, dim))
    want = np.array([O.min_distance64(cc, x) for x in q])
    np.testing.assert_allclose(validate.min_distance(cc, q), want, rtol=0, atol=1e-12)
    np.testing.assert_allclose(validate.min_distance(cc, q, cap=0.01), np.minimum(want, 0.01), rtol=0, atol=1e-12)


def test_validate_paths_accepts_oracle_paths_and_rejects_a_shortcut():
    sc = scenes.scene_c1()
    cc = sc.context()
    start, goals = query(sc, 1, 0)
    r = O.plan(O.Context(cc, 7), start, goals, num_new_samples=64, max_nodes=16384, min_vals=sc.robot.lower,
               max_vals=sc.robot.upper)
    assert r["status"] == O.STATUS_SOLVED
    path = r["path"]
    assert validate.validate_paths(cc, [path, None])[0]
    assert not validate.validate_paths(cc, [path, None])[1]
    # an obstacle placed on a robot sphere at an edge point between two waypoints is found
    ctx = O.Context(cc, 7)
    pts = validate.path_points(path, 16)
    mid = pts[len(pts) // 2 + 3]
    p = O.sphere_positions(ctx, mid)[10]
    moved = dict(cc, world_spheres=np.array([[p[0], p[1], p[2], 0.05]], np.float32))
    assert not validate.validate_paths(moved, [path])[0]


def test_host_loop_checker_matches_oracle_and_plans():
    sc = scenes.scene_c2()
    cc = sc.context()
    rng = np.random.default_rng(9)
    q = rng.uniform(sc.robot.lower, sc.robot.upper, size=(20_000, 7)).astype(np.float32)
    md = O.min_distance(O.Context(cc, 7, collision_mode=1, collision_margin=0.0), q)
    q = q[np.abs(md) > 1e-5][:3000]  # decidable in float32
    free = host_rrtc.BatchChecker(cc, 7).free(q)
    np.testing.assert_array_equal(free, ~O.in_collision(O.Context(cc, 7), q))
    sc1 = scenes.scene_c1()
    c1 = sc1.context()
    start, goals = query(sc1, 1, 1)
    r = host_rrtc.plan(host_rrtc.BatchChecker(c1, 7), start, goals, max_nodes=16384, min_vals=sc1.robot.lower,
                       max_vals=sc1.robot.upper, seed=1, time_limit=120.0)
    assert r["solved"] and not r["timed_out"]
    assert np.array_equal(r["path"][0], start) and np.array_equal(r["path"][-1], goals[0])
    assert validate.validate_paths(c1, [r["path"]], resolution=15)[0]


class _OracleEdges:
    """A CPU stand-in for PointCloudChecker.edges_free: the oracle's world verdict of every
    waypoint, 1 where a whole edge is free (the loop's host logic runs unchanged)."""

    def __init__(self, cc, dim):
        self.chk, self.dim, self.calls = host_rrtc.BatchChecker(cc, dim), dim, 0

    def edges_free(self, wp, world_geom=None, point_cloud=None, r_env=0.02):
        import torch

        self.calls += 1
        wp = np.asarray(wp, np.float32)
        return torch.from_numpy(self.chk.free(wp.reshape(-1, self.dim)).reshape(wp.shape[0], -1).all(1).astype(np.uint8))


class _NoEvent:
    def __init__(self, **_):
        pass

    def record(self):
        pass

    def elapsed_time(self, other):
        return 0.0


@pytest.mark.parametrize("kw", [dict(), dict(max_nodes=120), dict(connect_max_steps=3), dict(num_new_samples=8)])
def test_plan_robogpu_host_loop_equals_the_reference_loop(monkeypatch, kw):
    """prrtc_plan_robogpu's host logic (walk positions on the host, connect steps checked in
    blocks and replayed with array operations) equals oracle/host_rrtc.py's one-target-at-a-time
    restatement of the reference loop with the same checker: iterations, tree sizes, trees and
    path; also when the trees fill (max_nodes 120: query 0 fills both and never connects) and
    with short walks."""
    import torch

    from prrax_amd.robogpu import prrtc_plan_robogpu

    monkeypatch.setattr(torch.cuda, "Event", _NoEvent)
    sc = scenes.scene_c1()
    cc, robot = sc.context(), sc.robot
    for qi in range(3):
        start, goals = query(sc, 1, qi)
        args = dict(max_iterations=60, step_size=0.5, num_new_samples=32, granularity=16, max_nodes=4000,
                    connect_max_steps=32, seed=qi)
        args.update(kw)
        chk = _OracleEdges(cc, 7)
        res = prrtc_plan_robogpu(start, goals, robot, chk, **args)
        ref = host_rrtc.plan(host_rrtc.BatchChecker(cc, 7), start, goals, min_vals=robot.lower, max_vals=robot.upper,
                             **args)
        assert res.solved == ref["solved"] and res.iterations == ref["iterations"], (qi, kw)
        assert (res.tree_a_size, res.tree_b_size) == ref["sizes"], (qi, kw)
        if res.solved:
            np.testing.assert_array_equal(res.path, ref["path"])


def _c_layout() -> dict:
    src = r"""
    #include <stddef.h>
    #include <stdio.h>
    #include "prrtc_amd.h"
    int main(void) {
      printf("robot %zu\nworld %zu\nattrs %zu\noutputs %zu\nlaunch %zu\n", sizeof(prrtc_robot), sizeof(prrtc_world),
             sizeof(prrtc_attrs), sizeof(prrtc_outputs), sizeof(prrtc_launch));
      printf("world.coarse_sphere_link_idx %zu\nworld.n_self_pair_ranges %zu\nlaunch.fk_batch %zu\n",
             offsetof(prrtc_world, coarse_sphere_link_idx), offsetof(prrtc_world, n_self_pair_ranges),
             offsetof(prrtc_launch, fk_batch));
      return 0;
    }
    """
    return src


def test_ctypes_structs_match_the_header(tmp_path):
    c = tmp_path / "layout.c"
    c.write_text(_c_layout())
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(exe), str(c)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.splitlines())
    assert int(out["robot"]) == ctypes.sizeof(_lib.Robot)
    assert int(out["world"]) == ctypes.sizeof(_lib.World)
    assert int(out["attrs"]) == ctypes.sizeof(_lib.Attrs)
    assert int(out["outputs"]) == ctypes.sizeof(_lib.Outputs)
    assert int(out["launch"]) == ctypes.sizeof(_lib.Launch)
    assert int(out["world.coarse_sphere_link_idx"]) == _lib.World.coarse_sphere_link_idx.offset
    assert int(out["world.n_self_pair_ranges"]) == _lib.World.n_self_pair_ranges.offset
    assert int(out["launch.fk_batch"]) == _lib.Launch.fk_batch.offset


def test_coarse_keys_are_checked():
    cc = scenes.scene_c1().context()
    arr = check_context(cc, 7)
    assert arr["coarse_sphere_local"].shape == (0, 3) and arr["self_pair_ranges"].shape == (0, 2)
    bad = dict(cc, coarse_sphere_link_idx=np.zeros(3, np.int32), coarse_sphere_local=np.zeros((2, 3), np.float32),
               coarse_sphere_radius=np.zeros(3, np.float32))
    with pytest.raises(ValueError, match="coarse sphere count"):
        check_context(bad, 7)
    good = dict(bad, coarse_sphere_local=np.zeros((3, 3), np.float32))
    assert check_context(good, 7)["coarse_sphere_link_idx"].shape == (3,)
    assert oracle_free_fn(cc, 7)(np.zeros(7, np.float32)) in (True, False)
