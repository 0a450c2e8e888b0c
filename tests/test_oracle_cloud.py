"""The oracle's point-cloud verdict (robot sphere within r + r_env of a cloud point), pinned on
the CPU: exact boundary cases worked out by hand, negative radii, invalid spheres, and FP64 brute
force away from contact on the point-cloud scene."""

from __future__ import annotations

import numpy as np

from oracle import oracle as O
from prrax_amd import scenes

from test_oracle_fk_collision import _single_sphere_robot


def test_cloud_boundary_cases_by_hand():
    # robot sphere at the origin, r = 0.25; points on the x axis; r_env = 0.25: contact iff |x| <= 0.5
    cc = _single_sphere_robot({}, r=0.25)
    ctx = O.Context(cc, 1)
    z = np.zeros((1, 1), np.float32)
    for x, hit in ((0.5, True), (0.5000001, False), (-0.5, True), (0.25, True), (3.0, False)):
        assert bool(O.cloud_hits(ctx, z, np.array([[x, 0, 0]], np.float32), 0.25)[0]) == hit, x
    # (0.3, 0.4, 0): distance 0.5 exactly in float32 (0.09 + 0.16 rounds to 0.25)
    assert O.cloud_hits(ctx, z, np.array([[0.3, 0.4, 0.0]], np.float32), 0.25)[0]
    # a negative radius sum never touches, even at distance 0
    assert not O.cloud_hits(ctx, z, np.zeros((1, 3), np.float32), -0.3)[0]
    # no points: free
    assert not O.cloud_hits(ctx, z, np.zeros((0, 3), np.float32), 0.25)[0]


def test_cloud_skips_invalid_spheres():
    cc = _single_sphere_robot({}, r=0.25)
    cc["sphere_link_idx"] = np.array([-1], np.int32)
    ctx = O.Context(cc, 1)
    assert not O.cloud_hits(ctx, np.zeros((1, 1), np.float32), np.zeros((1, 3), np.float32), 0.25)[0]


def test_cloud_matches_fp64_away_from_contact():
    sc = scenes.scene_cloud(4000)
    pts, r_env = sc.extra["points"], sc.extra["r_env"]
    cc = sc.context()
    ctx = O.Context(cc, 7)
    rng = np.random.default_rng(3)
    q = rng.uniform(sc.robot.lower, sc.robot.upper, size=(300, 7)).astype(np.float32)
    got = O.cloud_hits(ctx, q, pts, r_env)
    r = cc["sphere_radius"].astype(np.float64)
    ok = cc["sphere_link_idx"] >= 0
    margin = []
    for qi in q.astype(np.float64):
        c = O.sphere_positions64(cc, qi)[ok]
        d = np.sqrt(((c[:, None, :] - pts[None].astype(np.float64)) ** 2).sum(-1)) - r[ok][:, None] - r_env
        margin.append(d.min())
    margin = np.asarray(margin)
    far = np.abs(margin) > 1e-5
    np.testing.assert_array_equal(got[far], margin[far] <= 0)
    assert far.mean() > 0.95 and 0.05 < got.mean() < 0.95


def test_host_cloud_checker_matches_oracle_away_from_contact():
    """The CPU baseline's k-d tree checker (oracle.host_rrtc.CloudChecker) agrees with the
    oracle's brute force wherever both the world and the cloud are decidable in float32."""
    from oracle import host_rrtc

    sc = scenes.scene_cloud(4000)
    pts, r_env = sc.extra["points"], sc.extra["r_env"]
    cc = sc.context()
    ctx = O.Context(cc, 7)
    rng = np.random.default_rng(8)
    q = rng.uniform(sc.robot.lower, sc.robot.upper, size=(400, 7)).astype(np.float32)
    want = ~(O.in_collision(ctx, q) | O.cloud_hits(ctx, q, pts, r_env))
    got = host_rrtc.CloudChecker(cc, 7, pts, r_env).free(q)
    world_md = np.abs(O.min_distance(O.Context(cc, 7, collision_mode=1, collision_margin=0.0), q))
    shrunk = O.cloud_hits(ctx, q, pts, r_env - 1e-5) == O.cloud_hits(ctx, q, pts, r_env + 1e-5)
    ok = (world_md > 1e-5) & shrunk
    assert ok.mean() > 0.9
    np.testing.assert_array_equal(got[ok], want[ok])
