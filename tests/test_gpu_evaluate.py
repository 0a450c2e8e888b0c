"""GPU: the benchmark-suite harness on a robot read from URDF/SRDF text (the synthetic 7-DOF arm)
with two problem sets of sphere + floor worlds; every solved path passes the FP64 re-check, and
the batched mode solves the same problems."""

from __future__ import annotations

import numpy as np
import pytest

from prrax_amd import robot_model as RM
from prrax_amd import scenes, validate
from prrax_amd.evaluate import evaluate_suite, summary

from test_robot_model import robot_urdf

pytestmark = pytest.mark.gpu

SRDF = '<robot name="arm"><disable_collisions link1="link1" link2="link3" reason="Never"/></robot>'


def _problems(robot, n_sets=2, per_set=6):
    base, lo, hi, _ = RM.load_robot(robot_urdf(robot), SRDF)
    out = {}
    for k in range(n_sets):
        rng = np.random.default_rng(100 + k)
        obstacles = [{"type": "sphere", "center": rng.uniform(scenes.REGION_LO, scenes.REGION_HI),
                      "radius": float(rng.uniform(0.05, 0.12))} for _ in range(30)]
        obstacles.append({"type": "halfspace", "normal": [0, 0, 1], "point": [0, 0, -0.2]})
        cc = dict(base, **RM.world_arrays(obstacles))
        free = lambda q: validate.min_distance(cc, q) > 1e-3  # noqa: E731
        items = []
        for i in range(per_set):
            sg = scenes.sample_free(rng, scenes.Robot(base, lo, hi), 3, free)
            items.append({"valid": i != 1, "start": sg[0].tolist(), "goals": sg[1:].tolist(), "obstacles": obstacles})
        out[f"set{k}"] = items
    return out


def test_suite_from_urdf_solves_and_validates():
    robot = scenes.make_arm(0, 48, 60)
    problems = _problems(robot)
    print("suite per query", flush=True)
    res = evaluate_suite(robot_urdf(robot), SRDF, problems, max_iterations=5000, num_new_samples=64,
                         max_nodes=65536)
    st = summary(res)
    assert (st["total"], st["valid"]) == (12, 10)
    assert st["solved"] == 10 and st["paths_invalid"] == 0
    print("suite batched", flush=True)
    resb = evaluate_suite(robot_urdf(robot), SRDF, problems, batched=True, max_iterations=5000, num_new_samples=64,
                          max_nodes=65536)
    assert resb.solved == res.solved and resb.path_invalid == 0
    assert [r["iterations"] for r in resb.rows] == [r["iterations"] for r in res.rows]
