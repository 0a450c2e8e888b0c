"""Pins the oracle planner by properties the reference kernel guarantees
(prrtc_planner.cu:675-1066) and by hand-checked tiny cases.

* Roots: start at index 0 of tree A (its own parent), goal g at index g of tree B.
* Every non-root node's parent is an earlier node of the same tree; the edge (parent -> node) is
  collision-free at the planner's waypoints t = (k+1)/G, rechecked with the FP64 checker away
  from contact.
* A solved path starts at the start, ends at a goal, and its consecutive nodes are tree edges
  (plus the near-zero junction between the two trees).
* Steps are at most step_size; the first round's samples are the Halton points of each
  explorer; iteration cap and tree overflow end the plan with their status.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from prrax_amd import scenes

from helpers import query

KW = dict(max_iterations=10_000, step_size=0.5, granularity=16, max_nodes=16384)


def _plan(sc, qi, S=64, n_goals=1, **kw):
    cc = sc.context()
    start, goals = query(sc, 1, qi, n_goals)
    args = dict(KW, num_new_samples=S, min_vals=sc.robot.lower, max_vals=sc.robot.upper)
    args.update(kw)
    return cc, start, goals, O.plan(O.Context(cc, sc.robot.dim), start, goals, **args), args


def _edges_free(cc, dim, tree, parents, size, roots, G, ctx):
    for i in range(size):
        p = int(parents[i])
        if i in roots:
            assert p == i
            continue
        assert 0 <= p < i, (i, p)
        a, b = tree[:, p], tree[:, i]
        for k in range(G):
            tf = np.float32(k + 1) / np.float32(G)
            q = (a + tf * (b - a)).astype(np.float32)
            assert not O.in_collision(ctx, q[None])[0]
            d = O.min_distance64(cc, q)
            assert d > -1e-5, (i, k, d)


@pytest.mark.parametrize("qi", [0, 1, 2])
def test_plan_properties(qi):
    sc = scenes.scene_c1()
    cc, start, goals, r, args = _plan(sc, qi, n_goals=2)
    assert r["status"] == O.STATUS_SOLVED
    sa, sb = r["sizes"]
    assert r["parents_a"][0] == 0 and np.array_equal(r["tree_a"][:, 0], start)
    for g in range(2):
        assert r["parents_b"][g] == g and np.array_equal(r["tree_b"][:, g], goals[g])
    ctx = O.Context(cc, 7)
    # edges of both trees (sampled: every node of small trees, otherwise 400 random nodes)
    rng = np.random.default_rng(qi)
    for tree, par, size, roots in ((r["tree_a"], r["parents_a"], sa, {0}), (r["tree_b"], r["parents_b"], sb, {0, 1})):
        idx = np.arange(size) if size <= 400 else np.sort(rng.choice(size, 400, replace=False))
        sub = {int(i) for i in idx}
        for i in sub:
            p = int(par[i])
            if i in roots:
                assert p == i
                continue
            assert 0 <= p < i
            step = np.linalg.norm(tree[:, i].astype(np.float64) - tree[:, p])
            assert step <= 0.5 + 1e-5
        _edges_free(cc, 7, tree, par, 0, roots, 16, ctx)
        for i in list(sub)[:60]:
            if i in roots:
                continue
            p = int(par[i])
            a, b = tree[:, p], tree[:, i]
            for k in range(16):
                tf = np.float32(k + 1) / np.float32(16)
                q = (a + tf * (b - a)).astype(np.float32)
                assert not O.in_collision(ctx, q[None])[0]
                assert O.min_distance64(cc, q) > -1e-5
    path = r["path"]
    assert np.array_equal(path[0], start)
    assert any(np.array_equal(path[-1], g) for g in goals)
    # consecutive path nodes: tree edges, except the junction where the connect walk ended
    jumps = np.linalg.norm(np.diff(path.astype(np.float64), axis=0), axis=1)
    assert np.all(jumps <= 0.5 + 1e-5)


def test_plan_is_deterministic():
    sc = scenes.scene_c1()
    _, _, _, r1, _ = _plan(sc, 3)
    _, _, _, r2, _ = _plan(sc, 3)
    for k in ("tree_a", "tree_b", "parents_a", "parents_b"):
        np.testing.assert_array_equal(r1[k], r2[k])
    assert r1["sizes"] == r2["sizes"] and r1["conn"] == r2["conn"]


def test_seed_changes_the_sample_sequence_only():
    sc = scenes.scene_c1()
    _, _, _, r0, _ = _plan(sc, 4, seed=0)
    _, _, _, r1, _ = _plan(sc, 4, seed=1)
    assert r0["status"] == r1["status"] == O.STATUS_SOLVED
    assert not (r0["sizes"] == r1["sizes"] and np.array_equal(r0["tree_a"], r1["tree_a"]))


def test_first_round_extends_toward_halton_samples():
    """With no obstacles, no dynamic domain and a huge step, round 1 appends each explorer's
    first Halton sample exactly (explorers < S/2 to tree A), in explorer order."""
    dim, S = 3, 8
    cc = {"fk_twists": np.zeros((0, 6), np.float32), "fk_parent_tf": np.zeros((0, 7), np.float32),
          "fk_parent_idx": np.zeros(0, np.int32), "fk_act_idx": np.zeros(0, np.int32),
          "fk_mimic_mul": np.zeros(0, np.float32), "fk_mimic_off": np.zeros(0, np.float32),
          "fk_mimic_act_idx": np.zeros(0, np.int32), "fk_topo_inv": np.zeros(0, np.int32),
          "sphere_link_idx": np.zeros(0, np.int32), "sphere_local": np.zeros((0, 3), np.float32),
          "sphere_radius": np.zeros(0, np.float32)}
    lo, hi = np.full(dim, -1, np.float32), np.full(dim, 1, np.float32)
    start = np.full(dim, -0.9, np.float32)
    goal = np.full((1, dim), 0.9, np.float32)
    r = O.plan(O.Context(cc, dim), start, goal, num_new_samples=S, step_size=100.0, dynamic_domain=False,
               min_vals=lo, max_vals=hi, max_nodes=64, max_iterations=1)
    for e in range(S):
        p = O.prime_order(e)
        sample = np.array([lo[d] + np.float32(O.halton(int(p[d]), 1)) * (hi[d] - lo[d]) for d in range(dim)],
                          np.float32)
        tree, idx, root = (r["tree_a"], 1 + e, start) if e < S // 2 else (r["tree_b"], 1 + e - S // 2, goal[0])
        # steer with scale 1 from the root: base + (sample - base) * 1 (prrtc_planner.cu:870-876)
        want = (root + (sample - root) * np.float32(1.0)).astype(np.float32)
        np.testing.assert_array_equal(tree[:, idx], want)
        assert np.allclose(want, sample, atol=1e-6)
    # explorer 0's connect walk reaches the goal tree in round 1 (free space)
    assert r["status"] == O.STATUS_SOLVED and r["iterations"] == 1


def test_iteration_cap_and_tree_full():
    sc = scenes.scene_c1()
    cc = sc.context()
    start, goals = query(sc, 1, 5)
    wall = dict(cc)
    wall["world_spheres"] = np.concatenate([cc["world_spheres"], np.array([[0, 0, 0, 10.0]], np.float32)])
    r = O.plan(O.Context(wall, 7), start, goals, num_new_samples=16, max_iterations=3, max_nodes=1000,
               min_vals=sc.robot.lower, max_vals=sc.robot.upper)
    assert r["status"] == O.STATUS_ITER_CAP and r["iterations"] == 4 and r["sizes"] == (1, 1)
    r = O.plan(O.Context(cc, 7), start, goals, num_new_samples=64, max_iterations=100, max_nodes=30,
               min_vals=sc.robot.lower, max_vals=sc.robot.upper)
    assert r["status"] == O.STATUS_TREE_FULL and max(r["sizes"]) == 30 and r["path"] is None


def test_dynamic_domain_radius_rules():
    """Radii: new nodes FLT_MAX; a collided extension turns FLT_MAX into dd_radius, then shrinks
    by (1 - alpha) down to dd_min_radius; a success grows it by (1 + alpha) (prrtc_planner.cu:46-83)."""
    sc = scenes.scene_c1()
    cc, start, goals, r, _ = _plan(sc, 6, S=128, dd_alpha=0.05, dd_radius=2.0, dd_min_radius=1.5)
    fmax = np.float32(np.finfo(np.float32).max)
    for radii, size in ((r["radii_a"], r["sizes"][0]), (r["radii_b"], r["sizes"][1])):
        rad = radii[:size]
        finite = rad[rad != fmax]
        assert np.all(finite >= np.float32(1.5) - 1e-6)
        assert np.all(finite <= np.float32(2.0) * np.float32(1.05) ** 200)


@pytest.mark.parametrize("bm", [0, 1, 2, 3])
def test_balance_modes_solve(bm):
    sc = scenes.scene_c1()
    _, _, _, r, _ = _plan(sc, 7, balance_mode=bm, tree_ratio=1.0 if bm == 2 else 0.5)
    assert r["status"] == O.STATUS_SOLVED


def test_trace_path_walks_to_roots():
    dim, N = 2, 8
    ta = np.zeros((dim, N), np.float32)
    tb = np.zeros((dim, N), np.float32)
    ta[:, :4] = np.array([[0, 1, 2, 3], [0, 0, 0, 0]], np.float32)
    tb[:, :3] = np.array([[9, 8, 7], [1, 1, 1]], np.float32)
    pa = np.array([0, 0, 1, 2, -1, -1, -1, -1], np.int32)
    pb = np.array([0, 0, 1, -1, -1, -1, -1, -1], np.int32)
    res = dict(tree_a=ta, tree_b=tb, parents_a=pa, parents_b=pb, sizes=(4, 3), conn=(3, 2, 0))
    path = O.trace_path(res)
    np.testing.assert_array_equal(path, [[0, 0], [1, 0], [2, 0], [3, 0], [7, 1], [8, 1], [9, 1]])
