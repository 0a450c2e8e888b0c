"""The drop-in boundary on the CPU: the C ABI library loads and exports every declared entry
point, and the Python API raises the reference's errors (prrtc.py:57-73, 346-381, 429-439)
before touching a device."""

from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import pytest

import prrax_amd
from prrax_amd import _lib, scenes
from prrax_amd.context import check_context

ROOT = Path(__file__).resolve().parent.parent


def declared_functions() -> list[str]:
    text = (ROOT / "include" / "prrtc_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(prrtc_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTED_SYMBOLS)
    assert b"gfx950" in lib.prrtc_version()


def test_library_built_from_these_sources():
    """prrtc_version() carries the hash of the sources the library was compiled from; the
    loader refuses a library whose hash differs from the sources next to it."""
    from prrax_amd import build

    have = _lib.library_source_hash(_lib.lib().prrtc_version().decode())
    assert have == build.source_hash()
    assert len(have) == 16
    with pytest.raises(RuntimeError, match="built from other sources"):
        _lib._check_fresh("prrtc_amd 0.3.0 (gfx950) src 0000000000000000")


def test_reference_api_surface():
    for name in ("prrtc_plan", "prrtc_plan_batch", "prrtc_nearest_neighbor", "prrtc_extend", "PRRTCResult"):
        assert hasattr(prrax_amd, name)
    fields = prrax_amd.PRRTCResult._fields
    assert fields[:11] == ("solved", "path", "tree_a_size", "tree_b_size", "iterations", "cost", "kernel_time_ms",
                           "tree_a_configs", "tree_b_configs", "tree_a_parents", "tree_b_parents")


def test_api_errors():
    from prrax_amd import prrtc_plan, prrtc_plan_batch

    cc = scenes.scene_c1().context()
    s = np.zeros(7, np.float32)
    g = np.zeros((1, 7), np.float32)
    with pytest.raises(ValueError, match="robogpu"):
        prrtc_plan(s, g, collision_context=cc, collision_checker="robogpu")
    with pytest.raises(ValueError, match="collision_checker must be one of"):
        prrtc_plan(s, g, collision_context=cc, collision_checker="nope")
    with pytest.raises(ValueError, match="single start"):
        prrtc_plan(np.zeros((2, 7)), g, collision_context=cc)
    with pytest.raises(ValueError, match="step_size"):
        prrtc_plan(s, g, collision_context=cc, step_size=0.0)
    with pytest.raises(ValueError, match="power of 2"):
        prrtc_plan(s, g, collision_context=cc, granularity=12)
    with pytest.raises(ValueError, match="<= 64"):
        prrtc_plan(s, g, collision_context=cc, granularity=128)
    with pytest.raises(ValueError, match="Collision-aware planning is the default"):
        prrtc_plan(s, g)
    bad = dict(cc)
    del bad["fk_twists"]
    with pytest.raises(ValueError, match="missing required keys"):
        prrtc_plan(s, g, collision_context=bad)
    with pytest.raises(ValueError, match="goal_configs must have shape"):
        prrtc_plan_batch(np.zeros((2, 7)), np.zeros((3, 1, 7)), collision_context=cc)
    with pytest.raises(ValueError, match="length"):
        prrtc_plan_batch(np.zeros((2, 7)), np.zeros((2, 1, 7)), collision_context=[cc])


def test_context_invariants():
    cc = scenes.scene_c1().context()
    check_context(cc, 7)
    for key, val, msg in [
        ("fk_act_idx", np.array([0, 1, 2, 3, 4, 5, 9], np.int32), "outside"),
        ("fk_mimic_act_idx", np.array([-1, -1, -1, -1, -1, -1, -2], np.int32), "outside"),
        ("fk_parent_idx", np.array([-1, 0, 1, 5, 3, 4, 5], np.int32), "parent"),
        ("fk_topo_inv", np.array([0, 1, 2, 3, 4, 5, 5], np.int32), "permutation"),
        ("sphere_link_idx", np.full(48, 7, np.int32), "sphere_link_idx"),
        ("world_spheres", np.zeros((3, 3), np.float32), r"\(N, 4\)"),
    ]:
        bad = dict(cc)
        bad[key] = val
        with pytest.raises(ValueError, match=msg):
            check_context(bad, 7)
    many = scenes.make_arm().ctx
    big = {k: np.concatenate([v] * 10) if k.startswith("fk_") else v for k, v in many.items()}
    big["fk_parent_idx"] = np.arange(-1, 69, dtype=np.int32)
    big["fk_topo_inv"] = np.arange(70, dtype=np.int32)
    with pytest.raises(ValueError, match="joint cap"):
        check_context(big, 7)


def test_default_queue_order_longest_first():
    """The default queue order of the one-work-group pass: a permutation, by start-to-nearest-
    goal distance times the square root of the obstacle load, descending; padded goal slots
    (beyond a query's goal count) are ignored."""
    from prrax_amd.prrtc import _default_order
    starts = np.zeros((4, 2), np.float32)
    goals = np.zeros((4, 2, 2), np.float32)
    goals[0, 0] = [1.0, 0.0]
    goals[1, 0] = [3.0, 0.0]
    goals[2, 0] = [2.0, 0.0]
    goals[3, 0] = [0.5, 0.0]
    goals[3, 1] = [0.0, 9.0]          # padding: query 3 has one goal, so this slot is ignored
    goals[:3, 1] = goals[:3, 0]
    gcounts = np.array([2, 2, 2, 1], np.int32)
    o = _default_order(starts, goals, gcounts, None)
    assert sorted(o.tolist()) == [0, 1, 2, 3]
    assert o.tolist() == [1, 2, 0, 3]
    # obstacle load: query 0 (distance 1) with 10 spheres outranks query 1 (distance 3), empty world
    def w(ns, nb=0, nc=0):
        return {"world_spheres": np.zeros((ns, 4)), "world_boxes": np.zeros((nb, 15)),
                "world_capsules": np.zeros((nc, 7))}
    o = _default_order(starts, goals, gcounts, [w(10), w(0), w(0, 1), w(0, 0, 1)])
    assert o.tolist() == [2, 0, 1, 3]  # keys 2 sqrt(5) = 4.47, sqrt(11) = 3.32, 3, 0.5 sqrt(3) = 0.87
