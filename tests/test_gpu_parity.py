"""GPU parity: the HIP library against the CPU oracle, bit for bit.

Every test calls the C ABI (through prrax_amd) on the device and compares with the oracle on
the same seeded inputs.  Integer/index results and every float produced by the shared
arithmetic (FK poses, tree configurations, distances) must be identical.
"""

from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

from oracle import oracle as O
from prrax_amd import scenes
from prrax_amd.prrtc import (CollisionChecker, forward_kinematics, prepare_batch, prrtc_extend,
                             prrtc_nearest_neighbor, prrtc_plan, prrtc_plan_batch)

from helpers import compare_plan, query

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _oracle_parallel(fn, qs, chunk=25_000):
    """fn over row chunks of qs on the host's cores (the oracle's ctypes calls release the GIL)."""
    parts = [qs[i:i + chunk] for i in range(0, qs.shape[0], chunk)]
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        return np.concatenate(list(ex.map(fn, parts)))


def _robot_scenes():
    return {"c1": scenes.scene_c1(), "c2": scenes.scene_c2(), "c4": scenes.scene_c4()}


@pytest.mark.parametrize("name", ["c1", "c2", "c4"])
def test_fk_bit_exact(name):
    sc = _robot_scenes()[name]
    cc = sc.context()
    dim = sc.robot.dim
    rng = np.random.default_rng(7)
    q = rng.uniform(-3.5, 3.5, size=(4096, dim)).astype(np.float32)
    say(f"fk {name} n=4096")
    got = forward_kinematics(cc, q).cpu().numpy()
    ctx = O.Context(cc, dim)
    want = np.stack([O.fk(ctx, qi) for qi in q])
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


# Minimum-distance values use the hardware square root (within 1 ulp); the distances in these
# scenes are below 4 m, where 1 ulp is at most 4.8e-7 m.
MIN_DIST_ATOL = 1e-6


@pytest.mark.parametrize("name,n", [("c1", 200_000), ("c2", 1_000_000), ("c4", 20_000)])
@pytest.mark.parametrize("mode", ["binary", "differentiable"])
def test_collision_bit_exact(name, n, mode):
    sc = _robot_scenes()[name]
    cc = sc.context()
    dim = sc.robot.dim
    rng = np.random.default_rng(11)
    q = rng.uniform(sc.robot.lower, sc.robot.upper, size=(n, dim)).astype(np.float32)
    say(f"collision {name} {mode} n={n}")
    # margin 0.005: the floor sits 0.01 m under the base-link spheres, so the reference's default
    # 0.02 margin would report every configuration in collision
    chk = CollisionChecker(cc, dim, collision_checker=mode, collision_margin=0.005)
    hit = chk.in_collision(q).cpu().numpy()
    mind = chk.min_distance(q[:5000]).cpu().numpy()
    ctx = O.Context(cc, dim, collision_mode=1 if mode == "differentiable" else 0, collision_margin=0.005)
    want = _oracle_parallel(lambda part: O.in_collision(ctx, part), q)
    np.testing.assert_array_equal(hit, want)
    np.testing.assert_allclose(mind, O.min_distance(ctx, q[:5000]), rtol=0, atol=MIN_DIST_ATOL)
    assert 0.05 < hit.mean() < 0.999  # both outcomes well represented


@pytest.mark.parametrize("mode", ["binary", "differentiable", "binary_coarse", "binary_coarse_coop"])
def test_collision_degenerate_geometry_bit_exact(mode):
    """Verdicts on worlds the coarse culling must not prune: negative radii (no contact when a
    radius sum is negative), non-orthonormal boxes, non-unit halfspace normals, degenerate capsules; a tree robot
    with prismatic, screw and mimic joints; configurations concentrated near contact."""
    from test_oracle_fk_collision import random_tree_robot

    cc, dim = random_tree_robot(3)
    rng = np.random.default_rng(21)
    ws = np.concatenate([rng.uniform(-0.6, 0.6, (60, 3)), rng.uniform(-0.05, 0.12, (60, 1))], 1)
    caps = np.concatenate([rng.uniform(-0.6, 0.6, (6, 6)), rng.uniform(-0.02, 0.05, (6, 1))], 1)
    caps[0, 3:6] = caps[0, :3]  # zero-length capsule
    boxes = np.zeros((6, 15))
    boxes[:, :3] = rng.uniform(-0.6, 0.6, (6, 3))
    for k in range(6):
        A = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        if k % 2:
            A = A * rng.uniform(0.5, 1.5, size=3) + 0.1 * rng.normal(size=(3, 3))  # skewed, not orthonormal
        boxes[k, 3:12] = A.T.reshape(-1)
        boxes[k, 12:] = rng.uniform(-0.02, 0.2, 3)
    hs = np.array([[0, 0, 1.3, 0, 0, -0.9], [0.2, 0.9, 0.1, 0, -1.2, 0]], np.float64)  # mostly on the free side
    cc.update(world_spheres=ws.astype(np.float32), world_capsules=caps.astype(np.float32),
              world_boxes=boxes.astype(np.float32), world_halfspaces=hs.astype(np.float32))
    q = rng.uniform(-2.8, 2.8, size=(60_000, dim)).astype(np.float32)
    flat = O.Context(cc, dim, collision_mode=1, collision_margin=0.0)
    md = O.min_distance(flat, q)
    q = q[np.argsort(np.abs(md))[:20_000]]  # the 20k configurations closest to contact
    say(f"degenerate geometry {mode}")
    chk = CollisionChecker(cc, dim, collision_checker=mode, collision_margin=0.01)
    hit = chk.in_collision(q).cpu().numpy()
    ctx = O.Context(cc, dim, collision_mode=1 if mode == "differentiable" else 0, collision_margin=0.01)
    np.testing.assert_array_equal(hit, O.in_collision(ctx, q))
    assert 0.05 < hit.mean() < 0.95


def test_nearest_neighbor_ties_and_parity():
    rng = np.random.default_rng(3)
    for dim, n, nq in ((7, 1000, 1), (7, 16384, 32), (14, 4096, 512), (3, 77, 40)):
        max_nodes = n + 13
        tree = np.zeros((dim, max_nodes), np.float32)
        tree[:, :n] = rng.uniform(-3, 3, size=(dim, n)).astype(np.float32)
        dup = rng.choice(n, size=n // 10, replace=False)
        tree[:, dup] = tree[:, dup[::-1]]  # duplicated nodes -> exact ties
        qs = rng.uniform(-3, 3, size=(nq, dim)).astype(np.float32)
        qs[: nq // 2] = tree[:, rng.choice(n, size=nq // 2)].T  # queries on nodes: distance 0 ties
        say(f"nn dim={dim} n={n} nq={nq}")
        dist, idx = prrtc_nearest_neighbor(tree, qs, n)
        oi, od = O.nearest(tree, n, qs)
        np.testing.assert_array_equal(idx.cpu().numpy(), oi)
        np.testing.assert_array_equal(dist.cpu().numpy().view(np.uint32), np.sqrt(od).view(np.uint32))


def test_extend_matches_steer():
    rng = np.random.default_rng(5)
    tree = rng.uniform(-3, 3, size=(7, 64)).astype(np.float32)
    nn = rng.integers(0, 64, size=100).astype(np.int32)
    smp = rng.uniform(-3, 3, size=(100, 7)).astype(np.float32)
    smp[0] = tree[:, nn[0]]  # zero-distance guard
    out, par, val = prrtc_extend(tree, nn, smp, 0.5)
    out = out.cpu().numpy()
    base = tree[:, nn].T
    d = np.sqrt(np.sum((smp - base).astype(np.float64) ** 2, axis=1))
    ref = base + (smp - base) * np.minimum(1.0, 0.5 / np.maximum(d, 1e-30))[:, None]
    np.testing.assert_allclose(out, ref, atol=2e-6)
    np.testing.assert_array_equal(par.cpu().numpy(), nn)
    assert (val.cpu().numpy() == 1).all()


C1_KW = dict(max_iterations=10_000, step_size=0.5, granularity=16, max_nodes=16384)


@pytest.mark.parametrize("qi,S", [(0, 128), (1, 128), (2, 32), (3, 64), (4, 16)])
def test_planner_c1_bit_exact(qi, S):
    sc = scenes.scene_c1()
    cc = sc.context()
    start, goals = query(sc, 1, qi)
    say(f"plan c1 q={qi} S={S}")
    res = prrtc_plan(start, goals, num_new_samples=S, min_vals=sc.robot.lower, max_vals=sc.robot.upper,
                     collision_context=cc, **C1_KW)
    ora = O.plan(O.Context(cc, 7), start, goals, num_new_samples=S, min_vals=sc.robot.lower,
                 max_vals=sc.robot.upper, **C1_KW)
    compare_plan(res, ora, f"c1 q{qi}")
    assert res.solved


@pytest.mark.parametrize("kw", [
    dict(balance_mode=0), dict(balance_mode=2, tree_ratio=1.0), dict(balance_mode=3), dict(dynamic_domain=False),
    dict(granularity=64), dict(granularity=1), dict(step_size=0.2), dict(collision_checker="differentiable"),
])
def test_planner_variants_bit_exact(kw):
    sc = scenes.scene_c1()
    cc = sc.context()
    start, goals = query(sc, 1, 10, n_goals=3)
    base = dict(C1_KW, num_new_samples=48, min_vals=sc.robot.lower, max_vals=sc.robot.upper)
    base.update(kw)
    say(f"plan variant {kw}")
    res = prrtc_plan(start, goals, collision_context=cc, **base)
    mode = 1 if base.pop("collision_checker", "binary") == "differentiable" else 0
    ora = O.plan(O.Context(cc, 7, collision_mode=mode, collision_margin=0.02), start, goals, **base)
    compare_plan(res, ora, str(kw))


@pytest.mark.parametrize("S,G,N,team", [(1, 1, 64, 0), (1, 16, 4096, 8), (3, 2, 200, 512), (130, 4, 3000, 2),
                                         (257, 16, 300, 64)])
def test_planner_edge_shapes_bit_exact(S, G, N, team):
    """Odd explorer counts (1, 3, one past a wave, one past a work-group), granularity 1/2/4,
    tiny trees and several team shapes, against the oracle."""
    sc = scenes.scene_c1()
    cc = sc.context()
    start, goals = query(sc, 1, 30, n_goals=2)
    kw = dict(num_new_samples=S, granularity=G, max_nodes=N, max_iterations=150, min_vals=sc.robot.lower,
              max_vals=sc.robot.upper)
    say(f"edge shape S={S} G={G} N={N} team={team}")
    res = prrtc_plan_batch(start[None], [goals], collision_context=cc, team_max=team, pass1_rounds=0, **kw)[0]
    ora = O.plan(O.Context(cc, 7), start, goals, **kw)
    compare_plan(res, ora, f"S{S} G{G} N{N}")


def test_planner_c2_bit_exact():
    """The large scene (64 spheres, 200 self pairs, 200 spheres, 16 boxes, 8 capsules, floor; 4
    goals, S = 512, max_nodes 65536): three seeded queries in one batch against the oracle."""
    sc = scenes.scene_c2()
    cc = sc.context()
    qs = [query(sc, 2, qi, n_goals=4) for qi in range(3)]
    kw = dict(num_new_samples=512, max_nodes=65536, max_iterations=10_000, min_vals=sc.robot.lower,
              max_vals=sc.robot.upper)
    say("plan c2 x3 batch vs oracle")
    res = prrtc_plan_batch(np.stack([s for s, _ in qs]), np.stack([g for _, g in qs]), collision_context=cc, **kw)
    with ThreadPoolExecutor(max_workers=3) as ex:
        oras = list(ex.map(lambda sg: O.plan(O.Context(cc, 7), sg[0], sg[1], **kw), qs))
    for qi, (r, ora) in enumerate(zip(res, oras)):
        compare_plan(r, ora, f"c2 q{qi}")
        assert r.solved


def test_planner_failure_modes_bit_exact():
    sc = scenes.scene_c1()
    cc = sc.context()
    start, goals = query(sc, 1, 20)
    common = dict(num_new_samples=64, min_vals=sc.robot.lower, max_vals=sc.robot.upper)
    # tree full
    say("plan tree-full")
    res = prrtc_plan(start, goals, collision_context=cc, max_nodes=16, max_iterations=100, **common)
    ora = O.plan(O.Context(cc, 7), start, goals, max_nodes=16, max_iterations=100, **common)
    compare_plan(res, ora, "tree-full")
    assert res.status == O.STATUS_TREE_FULL and not res.solved
    # iteration cap: an unreachable goal (in collision, isolated by a wall of spheres)
    say("plan iteration-cap")
    wall = dict(cc)
    wall["world_spheres"] = np.concatenate([cc["world_spheres"], np.array([[0.0, 0.0, 0.0, 10.0]], np.float32)])
    res = prrtc_plan(start, goals, collision_context=wall, max_nodes=4096, max_iterations=5, **common)
    ora = O.plan(O.Context(wall, 7), start, goals, max_nodes=4096, max_iterations=5, **common)
    compare_plan(res, ora, "iter-cap")
    assert res.status == O.STATUS_ITER_CAP and res.iterations == 6


def test_planner_no_collision_context():
    say("plan geometric")
    start = np.zeros(7, np.float32)
    goals = np.full((1, 7), 2.0, np.float32)
    res = prrtc_plan(start, goals, allow_unsafe_no_collision=True, max_nodes=4096, num_new_samples=32)
    ctx = O.Context(scenes.make_arm().ctx | {"sphere_link_idx": np.zeros(0, np.int32),
                                                "sphere_local": np.zeros((0, 3), np.float32),
                                                "sphere_radius": np.zeros(0, np.float32)}, 7)
    ora = O.plan(ctx, start, goals, max_nodes=4096, num_new_samples=32, max_iterations=1_000_000)
    compare_plan(res, ora, "geometric")
    assert res.solved


def test_batch_uneven_matches_single_queries():
    robot = scenes.make_arm(0, 64, 200)
    B = 12
    worlds, goals, starts, nodes, ctxs = [], [], [], [], []
    for qi in range(B):
        w, ng, N = scenes.c3_query_world(robot, qi)
        cc = dict(robot.ctx)
        cc.update(w)
        ctxs.append(cc)
        s, g = query(scenes.Scene(robot, w), 3, qi, n_goals=ng)
        starts.append(s)
        goals.append(g)
        nodes.append(N)
    kw = dict(num_new_samples=64, granularity=16, max_iterations=10_000, min_vals=robot.lower, max_vals=robot.upper)
    say(f"batch uneven B={B}")
    order = list(reversed(range(B)))
    res = prrtc_plan_batch(np.stack(starts), goals, collision_context=ctxs, max_nodes=max(nodes),
                           max_nodes_per_query=nodes, seeds=list(range(B)), order=order, **kw)
    for qi in range(B):
        ora = O.plan(O.Context(ctxs[qi], 7), starts[qi], goals[qi], max_nodes=nodes[qi], seed=qi, **kw)
        compare_plan(res[qi], ora, f"batch q{qi}")
    # the same queries one launch each give the same trees
    for qi in (0, 5):
        single = prrtc_plan(starts[qi], goals[qi], collision_context=ctxs[qi], max_nodes=nodes[qi], seed=qi, **kw)
        ora = O.plan(O.Context(ctxs[qi], 7), starts[qi], goals[qi], max_nodes=nodes[qi], seed=qi, **kw)
        compare_plan(single, ora, f"single q{qi}")


def test_dual_arm_batch_bit_exact():
    """14-DOF dual arm with mimic fingers (the dim > 8 kernels): a batch with uneven goal
    counts, step 0.3, against the oracle; then the same queries on teams (no pass 1)."""
    sc = scenes.scene_c4()
    cc = sc.context()
    starts, goals = [], []
    for qi in range(4):
        s, g = query(sc, 4, qi, n_goals=1 + qi % 4)
        starts.append(s)
        goals.append(g)
    kw = dict(step_size=0.3, num_new_samples=64, max_iterations=2000, max_nodes=16384, min_vals=sc.robot.lower,
              max_vals=sc.robot.upper)
    say("plan c4 batch")
    res = prrtc_plan_batch(np.stack(starts), goals, collision_context=cc, **kw)
    say("plan c4 batch on teams")
    res_t = prrtc_plan_batch(np.stack(starts), goals, collision_context=cc, pass1_rounds=0, team_max=16, **kw)
    for qi in range(4):
        ora = O.plan(O.Context(cc, 14), starts[qi], goals[qi], **kw)
        compare_plan(res[qi], ora, f"c4 q{qi}")
        compare_plan(res_t[qi], ora, f"c4 teams q{qi}")
        assert res[qi].solved


def test_graph_replay_matches_eager():
    """A batch captured in a HIP graph (reset kernel, pass 1, cooperative pass 2) replays to the
    eager result, with eager launches and synchronisations between the replays."""
    robot = scenes.make_arm(0, 64, 200)
    ctxs, starts, goals, nodes = [], [], [], []
    for qi in (0, 1, 2, 3, 26):
        w, ng, N = scenes.c3_query_world(robot, qi)
        cc = dict(robot.ctx)
        cc.update(w)
        s, g = query(scenes.Scene(robot, w), 3, qi, n_goals=ng)
        ctxs.append(cc); starts.append(s); goals.append(g); nodes.append(N)
    pb = prepare_batch(np.stack(starts), goals, collision_context=ctxs, max_nodes=max(nodes), max_nodes_per_query=nodes,
                       num_new_samples=64, max_iterations=300, min_vals=robot.lower, max_vals=robot.upper,
                       pass1_rounds=2)
    say("graph capture")
    pb.enqueue()
    torch.cuda.synchronize()
    ref = pb.bufs.results()
    ref_trees = [(r.tree_a_configs.cpu().numpy().copy(), r.tree_b_parents.cpu().numpy().copy()) for r in ref]
    g = pb.capture()
    for k in range(3):
        pb.bufs.ints.zero_()
        pb.bufs.tree_a.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = pb.bufs.results()
        for a, b, (ta, pbr) in zip(got, ref, ref_trees):
            assert a.status == b.status and a.iterations == b.iterations
            assert (a.tree_a_size, a.tree_b_size) == (b.tree_a_size, b.tree_b_size)
            np.testing.assert_array_equal(a.tree_a_configs.cpu().numpy(), ta)
            np.testing.assert_array_equal(a.tree_b_parents.cpu().numpy(), pbr)
        pb.enqueue()  # eager work between replays
        torch.cuda.synchronize()


def test_paths_valid_under_fp64_checker():
    """Every returned path's edges are free at 4x the planner's waypoint resolution under the
    FP64 checker (configs within 1e-5 m of contact are excluded as undecidable)."""
    sc = scenes.scene_c2()
    cc = sc.context()
    starts, goals = [], []
    for qi in range(6):
        s, g = query(sc, 2, qi, n_goals=4)
        starts.append(s)
        goals.append(g)
    say("plan c2 x6 batch")
    res = prrtc_plan_batch(np.stack(starts), np.stack(goals), collision_context=cc, num_new_samples=512,
                           max_nodes=65536, max_iterations=10_000, min_vals=sc.robot.lower, max_vals=sc.robot.upper)
    for i, r in enumerate(res):
        assert r.solved
        p = r.path.astype(np.float64)
        assert np.array_equal(r.path[0], starts[i])
        assert any(np.array_equal(r.path[-1], g) for g in goals[i])
        for a, b in zip(p[:-1], p[1:]):
            for t in np.linspace(0.0, 1.0, 17):  # the planner's waypoints t = k/16
                d = O.min_distance64(cc, a + t * (b - a))
                assert d > -1e-5, f"query {i}: path point in collision (d={d})"


@pytest.mark.parametrize("pass1_rounds,team_max,epoch_rounds,slots",
                         [(0, 1, 0, 0), (0, 2, 0, 0), (0, 8, 0, 0), (1, 4, 0, 0), (2, 32, 0, 0), (-1, 0, 0, 0),
                          (0, 512, 1, 0), (3, 0, 2, 0), (1, 16, 7, 0), (8, 0, 0, 0), (16, 0, 0, 0), (0, 0, 3, 300)])
def test_team_schedules_bit_exact(pass1_rounds, team_max, epoch_rounds, slots):
    """Hand-over budget, team size and epoch length change who computes what, never the result.
    Teams are slots / open queries work-groups (102 for 5 queries on 512 slots: not a power of
    two); a pass-1 budget above 8 rounds (the pass's minimum) exercises the early hand-over once
    every query has started; 300 slots give a grid barrier of 300 work-groups (arriving in 4
    groups)."""
    robot = scenes.make_arm(0, 64, 200)
    qids = [0, 1, 2, 26, 455]  # 26 needs 158 rounds, 455 never connects within the cap
    ctxs, starts, goals, nodes = [], [], [], []
    for qi in qids:
        w, ng, N = scenes.c3_query_world(robot, qi)
        cc = dict(robot.ctx)
        cc.update(w)
        s, g = query(scenes.Scene(robot, w), 3, qi, n_goals=ng)
        ctxs.append(cc); starts.append(s); goals.append(g); nodes.append(N)
    kw = dict(num_new_samples=128, granularity=16, max_iterations=40, min_vals=robot.lower, max_vals=robot.upper)
    say(f"team schedule pass1_rounds={pass1_rounds} team_max={team_max} epoch_rounds={epoch_rounds} slots={slots}")
    res = prrtc_plan_batch(np.stack(starts), goals, collision_context=ctxs, max_nodes=max(nodes),
                           max_nodes_per_query=nodes, pass1_rounds=pass1_rounds, team_max=team_max,
                           epoch_rounds=epoch_rounds, slots=slots, **kw)
    for k, qi in enumerate(qids):
        ora = _oracle_cache(qi, ctxs[k], starts[k], goals[k], nodes[k], kw)
        compare_plan(res[k], ora, f"c3 q{qi}")


_ORA = {}


def _oracle_cache(qi, cc, s, g, N, kw):
    if qi not in _ORA:
        _ORA[qi] = O.plan(O.Context(cc, 7), s, g, max_nodes=N, **kw)
    return _ORA[qi]
