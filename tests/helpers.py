"""Shared test helpers: scenes, oracle contexts and planner-result comparison."""

from __future__ import annotations

import numpy as np

from oracle import oracle as O
from prrax_amd import scenes


def oracle_free_fn(cc: dict, dim: int):
    ctx = O.Context(cc, dim, collision_mode=1, collision_margin=1e-5)
    return lambda q: O.min_distance(ctx, q) >= 1e-5


def query(scene: scenes.Scene, config: int, qi: int, n_goals: int = 1):
    cc = scene.context()
    rng = np.random.default_rng(scenes.query_seed(config, qi))
    sg = scenes.sample_free(rng, scene.robot, 1 + n_goals, oracle_free_fn(cc, scene.robot.dim))
    return sg[0], sg[1:]


def compare_plan(gpu, ora: dict, tag: str = "") -> None:
    """GPU PRRTCResult vs oracle plan dict, bit for bit."""
    assert gpu.status == ora["status"], f"{tag} status {gpu.status} != {ora['status']}"
    assert gpu.iterations == ora["iterations"], f"{tag} iterations {gpu.iterations} != {ora['iterations']}"
    assert (gpu.tree_a_size, gpu.tree_b_size) == tuple(ora["sizes"]), f"{tag} sizes"
    sa, sb = ora["sizes"]
    ta = gpu.tree_a_configs.cpu().numpy()
    tb = gpu.tree_b_configs.cpu().numpy()
    np.testing.assert_array_equal(ta.view(np.uint32), ora["tree_a"][:, :sa].view(np.uint32), err_msg=f"{tag} tree A")
    np.testing.assert_array_equal(tb.view(np.uint32), ora["tree_b"][:, :sb].view(np.uint32), err_msg=f"{tag} tree B")
    np.testing.assert_array_equal(gpu.tree_a_parents.cpu().numpy(), ora["parents_a"][:sa], err_msg=f"{tag} parents A")
    np.testing.assert_array_equal(gpu.tree_b_parents.cpu().numpy(), ora["parents_b"][:sb], err_msg=f"{tag} parents B")
    if ora["status"] == O.STATUS_SOLVED:
        np.testing.assert_array_equal(gpu.path.view(np.uint32), ora["path"].view(np.uint32), err_msg=f"{tag} path")
