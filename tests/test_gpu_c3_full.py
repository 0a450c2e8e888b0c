"""The benchmarked launch itself, at full size: the uneven 1024-query workload (7-DOF arm with 64
spheres and 200 self pairs; per query 10-300 world spheres, 0-24 boxes, 1-8 goals, max_nodes
8192-65536) planned in one batch at S = 512, G = 16, step 0.5, 10,000 iterations, with the
library's default launch (256 work-group slots, pass 1 then teams).

The oracle's plans of a sample of these queries (every 32nd, q26 and the two longest, q239 and
q455, which fill their trees) take it minutes, so they are stored as digests
(``tests/golden/make_c3_oracle.py``): each sampled query of the batch must equal its record bit
for bit (status, iterations, tree sizes, both trees' configurations and parents, the path).
Every other query is held to the properties of a plan: a solved path runs from its start to one
of its goals through tree nodes, and the batch solves at least 99% of the queries.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from prrax_amd import scenes
from prrax_amd.prrtc import CollisionChecker, prrtc_plan_batch

from helpers import plan_digests

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c3_s512_oracle.npz")


def say(msg):
    print(msg, flush=True)


def test_c3_1024_batch_matches_oracle_records():
    ref = np.load(FIXTURE)
    robot = scenes.make_arm(0, 64, 200)
    worlds, starts, goals, nodes = [], [], [], []
    say("c3 full: building 1024 queries")
    for qi in range(1024):
        w, ng, N = scenes.c3_query_world(robot, qi)
        cc = dict(robot.ctx)
        cc.update(w)
        free = CollisionChecker(cc, 7, collision_checker="differentiable", collision_margin=1e-5).is_free
        sg = scenes.sample_free(np.random.default_rng(scenes.query_seed(3, qi)), robot, 1 + ng, free)
        worlds.append(cc)
        starts.append(sg[0])
        goals.append(sg[1:])
        nodes.append(N)
    for k, qi in enumerate(ref["qids"]):
        ng = int(ref["n_goals"][k])
        assert nodes[qi] == ref["max_nodes"][k]
        np.testing.assert_array_equal(starts[qi].view(np.uint32), ref["starts"][k].view(np.uint32), err_msg=f"q{qi}")
        np.testing.assert_array_equal(goals[qi].view(np.uint32), ref["goals"][k, :ng].view(np.uint32), err_msg=f"q{qi}")
    say("c3 full: planning the batch (S 512)")
    res = prrtc_plan_batch(np.stack(starts), goals, collision_context=worlds, max_nodes=max(nodes),
                           max_nodes_per_query=nodes, max_iterations=10_000, step_size=0.5, num_new_samples=512,
                           granularity=16, min_vals=robot.lower, max_vals=robot.upper)
    for k, qi in enumerate(ref["qids"]):
        r = res[qi]
        tag = f"q{qi}"
        assert r.status == ref["status"][k], f"{tag} status {r.status} != {ref['status'][k]}"
        assert r.iterations == ref["iterations"][k], f"{tag} iterations {r.iterations} != {ref['iterations'][k]}"
        assert (r.tree_a_size, r.tree_b_size) == tuple(ref["sizes"][k]), f"{tag} sizes"
        got = plan_digests(r.tree_a_configs.cpu().numpy(), r.tree_b_configs.cpu().numpy(),
                           r.tree_a_parents.cpu().numpy(), r.tree_b_parents.cpu().numpy(), r.path)
        want = ref["digests"][k].reshape(5, 32)
        for name, g, w in zip(("tree A", "tree B", "parents A", "parents B", "path"), got, want):
            assert g == w.tobytes(), f"{tag} {name} differs from the oracle's"
    solved = 0
    for qi, r in enumerate(res):
        if not r.solved:
            continue
        solved += 1
        p = np.asarray(r.path)
        assert np.array_equal(p[0], starts[qi]), f"q{qi}: path does not start at the start"
        assert any(np.array_equal(p[-1], g) for g in goals[qi]), f"q{qi}: path does not end at a goal"
        ta = r.tree_a_configs.cpu().numpy().T
        tb = r.tree_b_configs.cpu().numpy().T
        nodes_ab = {row.tobytes() for row in ta} | {row.tobytes() for row in tb}
        assert all(row.tobytes() in nodes_ab for row in np.ascontiguousarray(p, np.float32)), \
            f"q{qi}: path point that is no tree node"
    say(f"c3 full: {solved}/1024 solved")
    assert solved >= 0.99 * 1024
