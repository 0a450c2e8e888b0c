"""Pins the oracle's planner rules with hand-computed cases, and shows that the cases bite.

Every GPU planner test compares the kernels with the oracle bit for bit, so a rule that the
oracle got wrong would be copied into the kernels unseen.  These cases pin each rule the GPU
inherits against values worked out by hand from the reference's code:

* dynamic domain (prrtc_planner.cu:46-83, 868): FLT_MAX -> dd_radius on the first collision,
  then x (1 - alpha) but not below dd_min_radius; x (1 + alpha) on success, FLT_MAX never grows;
  the skip test is strict (radius < distance);
* tree-side balance (prrtc_planner.cu:790-824): mode 0 halves, round 1 always mode 0, mode 1's
  near-balance split and ratio rule, mode 2's per-explorer alternation, the smaller-tree rule;
* steer (prrtc_planner.cu:870-876) and the connect walk's n = ceil(d / step) sub-edges of
  (to - from) / n (prrtc_planner.cu:975-993);
* the predicates' radius rules (prrtc_planner.cu:203-206, 250-304): no contact when r + R < 0;
* two planner runs small enough to follow by hand (one joint, one or two rounds).

The values below use constants that are exact in float32 (halves, quarters), so each expected
number is the rule's arithmetic done by hand, not the oracle's formula retyped.

``test_planted_mistake_is_caught`` then builds the oracle with each of a dozen deliberate
mistakes (``-DOC_PLANT=k`` in prrtc_oracle.c: a swapped comparison, a dropped term, floor for
ceil, ...) and checks that at least one case fails for every one of them.
"""

from __future__ import annotations

import ctypes
import subprocess

import numpy as np
import pytest

from oracle import build as oracle_build
from oracle import oracle as O

FLT_MAX = float(np.finfo(np.float32).max)

PLANTS = {
    1: "balance mode 1: ratio rule comparison swapped",
    2: "balance mode 1: near-balance split uses ratio instead of 1 - ratio",
    3: "balance mode 2: flips when the ratio is at or above tree_ratio",
    4: "dynamic domain: shrink multiplies by 1 + alpha",
    5: "dynamic domain: shrink clamps with min instead of max",
    6: "dynamic domain: grow applied to an unbounded radius",
    7: "dynamic domain: skip test not strict",
    8: "balance mode 0: explorer S/2 joins the start tree",
    9: "balance, smaller-tree rule: a tie goes to the goal tree",
    10: "steer: fraction d / step instead of step / d",
    11: "binary predicate: hit when the radius sum is negative",
    12: "connect walk: floor instead of ceil",
}


# ------------------------------------------------------------------------------ rule checks
def check_dynamic_domain(L=None):
    a, r0, rmin = 0.5, 4.0, 1.0
    assert O.dd_shrink(FLT_MAX, a, r0, rmin, L) == 4.0          # first collision: dd_radius
    assert O.dd_shrink(4.0, a, r0, rmin, L) == 2.0              # 4 (1 - 0.5)
    assert O.dd_shrink(2.0, a, r0, rmin, L) == 1.0
    assert O.dd_shrink(1.0, a, r0, rmin, L) == 1.0              # 0.5 clamped to dd_min_radius
    assert O.dd_shrink(1.5, a, r0, rmin, L) == 1.0              # 0.75 clamped
    assert O.dd_shrink(8.0, 0.25, r0, rmin, L) == 6.0
    assert O.dd_grow(4.0, a, L) == 6.0
    assert O.dd_grow(1.0, 0.25, L) == 1.25
    assert O.dd_grow(FLT_MAX, a, L) == FLT_MAX                   # unbounded stays unbounded
    # a whole sequence: collide, collide, succeed, succeed, collide x3, succeed
    r, seq = FLT_MAX, []
    for op in "ssggsssg":
        r = O.dd_shrink(r, a, r0, rmin, L) if op == "s" else O.dd_grow(r, a, L)
        seq.append(r)
    assert seq == [4.0, 2.0, 3.0, 4.5, 2.25, 1.125, 1.0, 1.5]
    # the library default alpha: FLT_MAX -> exactly dd_radius, then 4 (1 - 1e-4)
    assert O.dd_shrink(FLT_MAX, 1e-4, 4.0, 1.0, L) == 4.0
    assert abs(O.dd_shrink(4.0, 1e-4, 4.0, 1.0, L) - 3.9996) < 1e-6
    assert not O.dd_skips(FLT_MAX, 1e30, L)
    assert not O.dd_skips(4.0, 4.0, L)                          # strict: equal distance extends
    assert O.dd_skips(4.0, 4.5, L)
    assert not O.dd_skips(1.0, 0.5, L)


def check_balance(L=None):
    sides = lambda *a, **k: list(O.balance_sides(*a, L=L, **k))  # noqa: E731
    # mode 0: the first S/2 explorers extend the start tree
    assert sides(0, 5, 8, 3, 90, 0.5) == [0, 0, 0, 0, 1, 1, 1, 1]
    assert sides(0, 5, 7, 3, 90, 0.5) == [0, 0, 0, 1, 1, 1, 1]
    assert sides(0, 2, 1, 1, 1, 0.5) == [1]
    # round 1 uses the mode-0 split in every mode
    for mode in (1, 2, 3):
        assert sides(mode, 1, 8, 1, 50, 0.5) == [0, 0, 0, 0, 1, 1, 1, 1]
    # mode 1, |a - b| < int(1.5 S) = 12: explorers e < int(S (1 - a / (a + b))) extend A
    assert sides(1, 2, 8, 10, 12, 0.5) == [0] * 4 + [1] * 4      # 8 * 12/22 = 4.36 -> 4
    assert sides(1, 2, 8, 19, 8, 0.5) == [0] * 2 + [1] * 6       # 8 * 8/27 = 2.37 -> 2
    assert sides(1, 2, 8, 8, 19, 0.5) == [0] * 5 + [1] * 3       # 8 * 19/27 = 5.63 -> 5
    assert sides(1, 2, 8, 1, 1, 0.5) == [0] * 4 + [1] * 4        # 8 * 1/2 = 4
    # mode 1, |a - b| >= 12: every explorer extends A iff a / (a + b) < tree_ratio
    assert sides(1, 2, 8, 20, 8, 0.5) == [1] * 8                 # 20/28 = 0.71
    assert sides(1, 2, 8, 8, 20, 0.5) == [0] * 8                 # 8/28 = 0.29
    assert sides(1, 2, 8, 8, 20, 0.2) == [1] * 8                 # 0.29 >= 0.2
    assert sides(1, 2, 2, 1, 4, 0.5) == [0, 0]                   # diff 3 = int(1.5 * 2): ratio rule
    assert sides(1, 2, 2, 1, 3, 0.5) == [0, 1]                   # diff 2: 2 * 3/4 = 1.5 -> 1
    # mode 2: each explorer flips its tree when |t - o| / t < tree_ratio (t: its current tree)
    v = np.zeros(4, np.int32)
    assert sides(2, 2, 4, 10, 10, 0.5, vamp=v) == [1] * 4        # ratio 0: flip to B
    assert sides(2, 3, 4, 10, 10, 0.5, vamp=v) == [0] * 4        # and back
    assert sides(2, 2, 4, 100, 10, 0.5, vamp=np.zeros(4, np.int32)) == [0] * 4   # 90/100: stay on A
    assert sides(2, 2, 4, 100, 10, 0.5, vamp=np.ones(4, np.int32)) == [1] * 4    # 90/10: stay on B
    assert sides(2, 2, 4, 10, 14, 0.5, vamp=np.zeros(4, np.int32)) == [1] * 4    # 4/10 = 0.4: flip
    assert sides(2, 2, 4, 10, 14, 0.5, vamp=np.ones(4, np.int32)) == [0] * 4     # 4/14 = 0.29: flip
    v = np.array([0, 1, 0, 1], np.int32)
    assert sides(2, 2, 4, 10, 16, 0.5, vamp=v) == [0, 0, 0, 0]   # 6/10 = 0.6 stays; 6/16 = 0.375 flips
    assert list(v) == [0, 0, 0, 0]
    v = np.array([0, 1, 0, 1], np.int32)
    assert sides(2, 1, 4, 10, 16, 0.5, vamp=v) == [0, 0, 1, 1] and list(v) == [0, 1, 0, 1]  # round 1
    # any other mode: the smaller tree, ties to A
    assert sides(3, 2, 4, 5, 5, 0.5) == [0] * 4
    assert sides(3, 2, 4, 6, 5, 0.5) == [1] * 4
    assert sides(3, 2, 4, 4, 5, 0.5) == [0] * 4


def check_steer(L=None):
    np.testing.assert_array_equal(O.steer([1, 1], [5, 1], 4.0, 1.0, L), [2, 1])        # step / d = 1/4
    np.testing.assert_array_equal(O.steer([0, 2], [0, -6], 8.0, 2.0, L), [0, 0])       # 1/4 again
    np.testing.assert_array_equal(O.steer([1, 1], [1.5, 1], 0.5, 1.0, L), [1.5, 1])    # d < step: the sample
    np.testing.assert_array_equal(O.steer([2, 2], [2, 2], 0.0, 0.5, L), [2, 2])        # d = 0
    np.testing.assert_array_equal(O.steer([0.0], [1e-9], 1e-9, 0.5, L), np.float32([1e-9]))  # d <= 1e-8


def check_connect(L=None):
    n, v = O.connect_steps([0.0], [1.0], 1.0, 0.5, L)
    assert n == 2 and v[0] == 0.5
    n, v = O.connect_steps([0.0, 0.0], [0.75, -1.5], 1.5, 0.5, L)    # 1.5 / 0.5 = 3 exactly
    assert n == 3 and list(v) == [0.25, -0.5]
    n, v = O.connect_steps([0.0, 0.0], [0.3, 0.4], 0.5, 0.5, L)
    assert n == 1 and np.array_equal(v, np.float32([0.3, 0.4]))
    n, v = O.connect_steps([1.0], [2.0], 0.3, 0.5, L)                 # ceil(0.6) = 1
    assert n == 1 and v[0] == 1.0
    d = float(np.nextafter(np.float32(1.0), np.float32(2.0)))         # 1 + 2^-23: 2.0000002 steps
    n, v = O.connect_steps([0.0], [d], d, 0.5, L)
    assert n == 3
    for dist, step in ((0.0, 0.5), (5e-9, 0.5), (1.0, 5e-9)):
        n, v = O.connect_steps([0.0], [1.0], dist, step, L)
        assert n == 0 and v[0] == 0.0


def _slider_context(obstacles, radius=0.1, world_key="world_spheres"):
    """A one-joint robot sliding along x (prismatic twist (1,0,0,0,0,0)) carrying one sphere of
    the given radius at its joint origin: the sphere's centre is (q, 0, 0)."""
    cc = {
        "fk_twists": np.array([[1, 0, 0, 0, 0, 0]], np.float32),
        "fk_parent_tf": np.array([[1, 0, 0, 0, 0, 0, 0]], np.float32),
        "fk_parent_idx": np.array([-1], np.int32), "fk_act_idx": np.array([0], np.int32),
        "fk_mimic_mul": np.array([1], np.float32), "fk_mimic_off": np.array([0], np.float32),
        "fk_mimic_act_idx": np.array([-1], np.int32), "fk_topo_inv": np.array([0], np.int32),
        "sphere_link_idx": np.array([0], np.int32), "sphere_local": np.zeros((1, 3), np.float32),
        "sphere_radius": np.array([radius], np.float32),
    }
    cc[world_key] = np.asarray(obstacles, np.float32)
    return cc


def check_predicates(L=None):
    # sphere of radius 0.05 at x = q against a world sphere at the origin
    def hit(q, R, r=0.05, mode=0, margin=0.0):
        ctx = O.Context(_slider_context([[0, 0, 0, R]], radius=r), 1, collision_mode=mode, collision_margin=margin)
        return bool(O.in_collision(ctx, [[q]], L=L)[0])

    assert hit(0.125, 0.1) and not hit(0.25, 0.1)                 # contact at 0.15
    assert not hit(0.01, -0.1) and not hit(0.0, -0.06)              # r + R < 0: never in contact
    assert hit(0.0, -0.05)                                          # r + R = 0 at distance 0
    assert hit(0.25, 0.1, mode=1, margin=0.125) and not hit(0.5, 0.1, mode=1, margin=0.25)
    assert not hit(0.25, -0.5, mode=1, margin=0.125)                # L = -0.325 < 0
    # box with half extents 0.5 at the origin: a negative sphere radius needs the centre deeper
    # than |r| inside
    box = [[0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0.5, 0.5, 0.5]]

    def hit_box(q, r, mode=0, margin=0.0):
        ctx = O.Context(_slider_context(box, radius=r, world_key="world_boxes"), 1, collision_mode=mode,
                        collision_margin=margin)
        return bool(O.in_collision(ctx, [[q]], L=L)[0])

    assert hit_box(0.625, 0.25) and not hit_box(0.875, 0.25)
    assert hit_box(0.25, -0.125) and not hit_box(0.4375, -0.125) and not hit_box(0.75, -0.125)
    assert hit_box(0.75, 0.0, mode=1, margin=0.5) and not hit_box(1.25, 0.0, mode=1, margin=0.5)
    assert hit_box(0.25, -0.125, mode=1, margin=0.0) and not hit_box(0.4375, -0.125, mode=1, margin=0.0)


def check_tiny_plans(L=None):
    """Two runs small enough to follow by hand: a one-joint slider (q in [-1, 2], step 0.5,
    one explorer, G = 16), so explorer 0's first samples are -1 + 3 h with h = 1/3, 2/3, i.e.
    exactly 0.0 and 1.0."""
    kw = dict(num_new_samples=1, granularity=16, step_size=0.5, max_nodes=64, min_vals=[-1.0], max_vals=[2.0],
              balance_mode=1, tree_ratio=0.5)
    # (a) no obstacle.  Round 1: S/2 = 0, so the explorer extends tree B (the goal 1.0) toward
    # the sample 0.0: d = 1, steer 1 + (0 - 1) * 0.5 = 0.5 (B[1], parent 0).  Connect to tree A
    # (start 0.0): d = 0.5, ceil(0.5 / 0.5) = 1 sub-edge of -0.5, appended as B[2] = 0.0 (parent 1);
    # the walk reached A[0]: solved with connection (a 0, b 2, tree B).
    free = _slider_context(np.zeros((0, 4), np.float32))
    free["sphere_radius"] = np.zeros(0, np.float32)
    free["sphere_link_idx"] = np.zeros(0, np.int32)
    free["sphere_local"] = np.zeros((0, 3), np.float32)
    r = O.plan(O.Context(free, 1), [0.0], [[1.0]], max_iterations=10, L=L, **kw)
    assert r["status"] == O.STATUS_SOLVED and r["iterations"] == 1
    assert r["sizes"] == (1, 3) and r["conn"] == (0, 2, 1)
    assert list(r["tree_b"][0, :3]) == [1.0, 0.5, 0.0] and list(r["parents_b"][:3]) == [0, 0, 1]
    assert list(r["path"][:, 0]) == [0.0, 0.0, 0.5, 1.0]
    # (b) a sphere of radius 0.1 on the slider, a world sphere (0.6, 0, 0, 0.1): q in [0.4, 0.8]
    # collides, so start 0.0 and goal 1.0 never connect.  alpha = 0.5, dd_radius 4, min 1.
    #   round 1 (mode-0 split: B): 1.0 -> sample 0.0, steer to 0.5, the waypoint 0.5 collides:
    #     B[0]'s radius FLT_MAX -> 4.
    #   round 2 (sizes 1, 1: |0| < int(1.5), e < int(1 * 0.5) = 0 is false: B): sample 1.0 is
    #     at distance 0 of B[0] (4 < 0 is false: no skip), steer keeps 1.0, the edge is free:
    #     B[1] = 1.0 (parent 0), B[0]'s radius 4 -> 6.  Connect to A[0] = 0.0: d = 1, two
    #     sub-edges of -0.5; the first reaches 0.5 and collides, so nothing is appended.
    #   round 3 exceeds max_iterations = 2: iteration cap with iterations = 3.
    cc = _slider_context([[0.6, 0, 0, 0.1]])
    r = O.plan(O.Context(cc, 1), [0.0], [[1.0]], max_iterations=2, dd_alpha=0.5, dd_radius=4.0, dd_min_radius=1.0,
               L=L, **kw)
    assert r["status"] == O.STATUS_ITER_CAP and r["iterations"] == 3
    assert r["sizes"] == (1, 2)
    assert list(r["tree_b"][0, :2]) == [1.0, 1.0] and list(r["parents_b"][:2]) == [0, 0]
    assert r["radii_b"][0] == 6.0 and r["radii_b"][1] == FLT_MAX and r["radii_a"][0] == FLT_MAX


CHECKS = [check_dynamic_domain, check_balance, check_steer, check_connect, check_predicates, check_tiny_plans]


@pytest.mark.parametrize("check", CHECKS, ids=lambda c: c.__name__)
def test_rule(check):
    check()


@pytest.fixture(scope="module")
def planted_libs(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_plants")
    libs = {}
    for k in PLANTS:
        out = d / f"liboracle_plant{k}.so"
        subprocess.run(["gcc", *oracle_build.FLAGS, f"-DOC_PLANT={k}", "-o", str(out), str(oracle_build.SRC), "-lm"],
                       check=True)
        libs[k] = O.bind(ctypes.CDLL(str(out)))
    return libs


@pytest.mark.parametrize("k", sorted(PLANTS))
def test_planted_mistake_is_caught(planted_libs, k):
    caught = []
    for check in CHECKS:
        try:
            check(planted_libs[k])
        except AssertionError:
            caught.append(check.__name__)
    assert caught, f"planted mistake {k} ({PLANTS[k]}) passed every rule check"
