"""The stored oracle records of the full-size 1024-query workload (``tests/golden``) still match
the oracle: two short sampled queries are re-planned and compared with their records, so an
oracle change that leaves the records stale fails here before the GPU test reads them."""

from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import oracle as O
from prrax_amd import scenes

from helpers import plan_digests, query

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c3_s512_oracle.npz")


@pytest.mark.parametrize("qi", [32, 256])
def test_c3_record_matches_the_oracle(qi):
    ref = np.load(FIXTURE)
    k = int(np.flatnonzero(ref["qids"] == qi)[0])
    robot = scenes.make_arm(0, 64, 200)
    w, ng, N = scenes.c3_query_world(robot, qi)
    cc = dict(robot.ctx)
    cc.update(w)
    s, g = query(scenes.Scene(robot, w), 3, qi, n_goals=ng)
    assert N == ref["max_nodes"][k] and ng == ref["n_goals"][k]
    np.testing.assert_array_equal(s, ref["starts"][k])
    np.testing.assert_array_equal(g, ref["goals"][k, :ng])
    r = O.plan(O.Context(cc, 7), s, g, max_nodes=N, max_iterations=10_000, step_size=0.5, num_new_samples=512,
               granularity=16, min_vals=robot.lower, max_vals=robot.upper)
    sa, sb = r["sizes"]
    assert (r["status"], r["iterations"], sa, sb) == (ref["status"][k], ref["iterations"][k], *ref["sizes"][k])
    d = plan_digests(r["tree_a"][:, :sa], r["tree_b"][:, :sb], r["parents_a"][:sa], r["parents_b"][:sb], r["path"])
    assert b"".join(d) == ref["digests"][k].tobytes()
