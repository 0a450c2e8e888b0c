"""Pins the oracle's FK and collision predicates against the FP64 restatement and closed forms.

The FP64 side (``oracle.fk64`` etc.) is an independent formulation: 4x4 homogeneous matrices,
Rodrigues' formula (itself checked against a matrix exponential), and textbook point-segment,
point-box and point-plane distances.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import oracle as O
from prrax_amd import scenes


def quat_to_mat(q):
    return O.quat_mat64(q)


def random_tree_robot(seed: int) -> tuple[dict, int]:
    """A tree-shaped robot with fixed, revolute, prismatic, screw and mimic joints, stored in a
    permuted (non-identity topo_inv) order."""
    rng = np.random.default_rng(seed)
    nj, dim = 10, 6
    parent = np.array([-1, 0, 1, 1, 3, 4, 0, 6, 7, 2], np.int32)
    kinds = ["rev", "rev", "pri", "rev", "screw", "fixed", "rev", "rev", "mimic", "rev"]
    tw = np.zeros((nj, 6), np.float32)
    act = np.full(nj, -1, np.int32)
    mim = np.full(nj, -1, np.int32)
    mul = np.ones(nj, np.float32)
    off = np.zeros(nj, np.float32)
    a = 0
    for j, k in enumerate(kinds):
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        if k == "rev":
            tw[j, 3:] = ax
        elif k == "pri":
            tw[j, :3] = ax
        elif k == "screw":
            tw[j, 3:] = ax
            tw[j, :3] = rng.normal(size=3) * 0.3
        elif k == "mimic":
            tw[j, 3:] = ax
        if k in ("rev", "pri", "screw"):
            act[j] = a % dim
            a += 1
        if k == "mimic":
            mim[j] = 2
            mul[j] = -0.7
            off[j] = 0.3
    ptf = np.zeros((nj, 7), np.float32)
    qs = rng.normal(size=(nj, 4))
    qs /= np.linalg.norm(qs, axis=1, keepdims=True)
    ptf[:, :4] = qs
    ptf[:, 4:] = rng.uniform(-0.3, 0.3, size=(nj, 3))
    topo_inv = rng.permutation(nj).astype(np.int32)
    ns = 24
    cc = {
        "fk_twists": tw, "fk_parent_tf": ptf, "fk_parent_idx": parent, "fk_act_idx": act, "fk_mimic_mul": mul,
        "fk_mimic_off": off, "fk_mimic_act_idx": mim, "fk_topo_inv": topo_inv,
        "sphere_link_idx": np.concatenate([rng.integers(0, nj, size=ns - 2), [-1, 3]]).astype(np.int32),
        "sphere_local": rng.uniform(-0.1, 0.1, size=(ns, 3)).astype(np.float32),
        "sphere_radius": rng.uniform(0.02, 0.06, size=ns).astype(np.float32),
        "self_pairs": np.array([[0, 5], [1, 7], [2, 20], [3, 22], [30, 1], [-1, 2]], np.int32),
    }
    return cc, dim


def _robots():
    out = {"c1": (scenes.scene_c1().context(), 7), "c2": (scenes.scene_c2().context(), 7),
           "c4": (scenes.scene_c4().context(), 14)}
    out["tree"] = random_tree_robot(3)
    return out


def test_rodrigues_matches_matrix_exponential():
    rng = np.random.default_rng(0)
    for _ in range(20):
        tw = rng.normal(size=6)
        th = rng.uniform(-3, 3)
        xi = np.zeros((4, 4))
        w, v = tw[3:] * th, tw[:3] * th
        xi[:3, :3] = O._skew(w)
        xi[:3, 3] = v
        np.testing.assert_allclose(O.se3_exp64(tw, th), expm(xi), atol=1e-10)


@pytest.mark.parametrize("name", ["c1", "c2", "c4", "tree"])
def test_fk_fp32_within_1e5_of_fp64(name):
    cc, dim = _robots()[name]
    ctx = O.Context(cc, dim)
    rng = np.random.default_rng(1)
    worst_p, worst_r = 0.0, 0.0
    for _ in range(300):
        q = rng.uniform(-3.2, 3.2, size=dim).astype(np.float32)
        T32 = O.fk(ctx, q)
        T64 = O.fk64(cc, q)
        for j in range(T32.shape[0]):
            worst_p = max(worst_p, float(np.max(np.abs(T32[j, 4:] - T64[j, :3, 3]))))
            worst_r = max(worst_r, float(np.max(np.abs(quat_to_mat(T32[j, :4]) - T64[j, :3, :3]))))
        n = np.linalg.norm(T32[:, :4], axis=1)
        assert np.all(np.abs(n - 1) < 1e-5)
    assert worst_p < 1e-5 and worst_r < 1e-5, (worst_p, worst_r)


def test_fk_detects_a_transposed_rotation():
    """The FK pin fails for a plausible mistake: conjugated (transposed) joint rotations."""
    cc, dim = _robots()["c1"]
    ctx = O.Context(cc, dim)
    q = np.full(dim, 0.7, np.float32)
    T32 = O.fk(ctx, q)
    wrong = T32.copy()
    wrong[:, 1:4] *= -1
    T64 = O.fk64(cc, q)
    err = max(float(np.max(np.abs(quat_to_mat(wrong[j, :4]) - T64[j, :3, :3]))) for j in range(T32.shape[0]))
    assert err > 1e-2


def test_fk_zero_config_is_product_of_parent_transforms():
    cc, dim = _robots()["c1"]
    T = O.fk(O.Context(cc, dim), np.zeros(dim, np.float32))
    acc = np.zeros(3)
    for j in range(T.shape[0]):
        acc += cc["fk_parent_tf"][j, 4:]
        np.testing.assert_allclose(T[j, 4:], acc, atol=1e-6)
        np.testing.assert_allclose(T[j, :4], [1, 0, 0, 0], atol=1e-7)


def test_fk_single_revolute_is_rodrigues():
    ax = np.array([0.3, -0.4, np.sqrt(1 - 0.25)], np.float32)
    cc = {"fk_twists": np.array([[0, 0, 0, *ax]], np.float32), "fk_parent_tf": np.array([[1, 0, 0, 0, 0.1, 0.2, 0.3]], np.float32),
          "fk_parent_idx": np.array([-1], np.int32), "fk_act_idx": np.array([0], np.int32),
          "fk_mimic_mul": np.ones(1, np.float32), "fk_mimic_off": np.zeros(1, np.float32),
          "fk_mimic_act_idx": np.array([-1], np.int32), "fk_topo_inv": np.array([0], np.int32),
          "sphere_link_idx": np.zeros(0, np.int32), "sphere_local": np.zeros((0, 3), np.float32),
          "sphere_radius": np.zeros(0, np.float32)}
    ctx = O.Context(cc, 1)
    for th in (-2.5, -0.3, 0.0, 1.0, 3.0):
        T = O.fk(ctx, np.array([th], np.float32))[0]
        a = ax.astype(np.float64) / np.linalg.norm(ax)
        K = O._skew(a)
        R = np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * K @ K
        np.testing.assert_allclose(quat_to_mat(T[:4]), R, atol=2e-6)
        np.testing.assert_allclose(T[4:], [0.1, 0.2, 0.3], atol=1e-7)


def test_fk_mimic_equals_driving_joint():
    cc, dim = _robots()["c4"]
    ctx = O.Context(cc, dim)
    q = np.random.default_rng(2).uniform(-2, 2, size=dim).astype(np.float32)
    T = O.fk(ctx, q)
    # finger joints 7, 8 of arm 0 mimic actuated joint 6 (mul 1, off 0): their rotation about x
    # relative to the wrist frame equals q[6]
    for f in (7, 8):
        Rw = quat_to_mat(T[6, :4])
        Rf = quat_to_mat(T[f, :4])
        rel = Rw.T @ Rf
        ang = math.atan2(rel[2, 1], rel[1, 1])
        assert abs(ang - q[6]) < 2e-6


# ----------------------------------------------------------------------------- collision
@pytest.mark.parametrize("name", ["c1", "c2", "c4", "tree"])
def test_min_distance_matches_fp64(name):
    cc, dim = _robots()[name]
    if name == "tree":
        cc = dict(cc, **scenes.make_world(np.random.default_rng(4), 30, 0.05, 0.15, n_boxes=4, n_capsules=4,
                                          floor=np.array([[0, 0, 1, 0, 0, -0.5]], np.float32)))
    ctx = O.Context(cc, dim)
    rng = np.random.default_rng(5)
    q = rng.uniform(-3, 3, size=(200, dim)).astype(np.float32)
    md32 = O.min_distance(ctx, q)
    md64 = np.array([O.min_distance64(cc, qi) for qi in q])
    np.testing.assert_allclose(md32, md64, atol=1e-5)


@pytest.mark.parametrize("name", ["c1", "c2", "c4"])
def test_binary_predicate_matches_fp64_away_from_contact(name):
    cc, dim = _robots()[name]
    ctx = O.Context(cc, dim)
    rng = np.random.default_rng(6)
    q = rng.uniform(-2.8, 2.8, size=(3000 if name != "c4" else 600, dim)).astype(np.float32)
    hit = O.in_collision(ctx, q)
    md64 = np.array([O.min_distance64(cc, qi) for qi in q])
    decided = np.abs(md64) > 1e-5
    assert decided.mean() > 0.99
    np.testing.assert_array_equal(hit[decided], (md64 <= 0)[decided])
    assert 0.05 < hit.mean() < 0.999


def test_margin_predicate_is_min_distance_below_margin():
    cc, dim = _robots()["c1"]
    ctx = O.Context(cc, dim, collision_mode=1, collision_margin=0.02)
    q = np.random.default_rng(8).uniform(-2.8, 2.8, size=(2000, dim)).astype(np.float32)
    np.testing.assert_array_equal(O.in_collision(ctx, q), O.min_distance(ctx, q) < 0.02)


def _single_sphere_robot(world: dict, r=0.1):
    cc = {"fk_twists": np.zeros((1, 6), np.float32), "fk_parent_tf": np.array([[1, 0, 0, 0, 0, 0, 0]], np.float32),
          "fk_parent_idx": np.array([-1], np.int32), "fk_act_idx": np.array([-1], np.int32),
          "fk_mimic_mul": np.ones(1, np.float32), "fk_mimic_off": np.zeros(1, np.float32),
          "fk_mimic_act_idx": np.array([-1], np.int32), "fk_topo_inv": np.array([0], np.int32),
          "sphere_link_idx": np.array([0], np.int32), "sphere_local": np.zeros((1, 3), np.float32),
          "sphere_radius": np.array([r], np.float32)}
    cc.update(world)
    return cc


def test_primitive_closed_forms():
    # sphere at the origin with r = 0.1 (fixed joint), one primitive at a time
    z = np.zeros(1, np.float32)
    cases = [
        ({"world_spheres": np.array([[0.35, 0, 0, 0.2]], np.float32)}, 0.05),
        ({"world_capsules": np.array([[1, 1, 0, 1, -1, 0, 0.1]], np.float32)}, 0.8),   # segment x=1
        ({"world_capsules": np.array([[0.5, 0, 0, 2, 0, 0, 0.1]], np.float32)}, 0.3),  # endpoint closest
        ({"world_boxes": np.array([[1, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0.5, 0.5, 0.5]], np.float32)}, 0.4),
        ({"world_boxes": np.array([[1, 1, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0.5, 0.5, 0.5]], np.float32)},
         math.sqrt(0.5) - 0.1),                                                       # corner region
        ({"world_boxes": np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0.5, 0.3, 0.2]], np.float32)}, -0.3),  # inside
        ({"world_halfspaces": np.array([[0, 0, 1, 0, 0, -0.5]], np.float32)}, 0.4),
    ]
    for world, want in cases:
        cc = _single_sphere_robot(world)
        ctx = O.Context(cc, 1)
        got = float(O.min_distance(ctx, z)[0])
        assert abs(got - want) < 1e-6, (world, got, want)
        assert abs(O.min_distance64(cc, z) - want) < 1e-7  # inputs are float32-rounded
        assert bool(O.in_collision(ctx, z)[0]) == (want <= 0)


def test_invalid_spheres_and_pairs_are_skipped():
    cc, dim = _robots()["tree"]
    ctx = O.Context(cc, dim)
    # sphere 22 has link -1, pairs with out-of-range or invalid-link members are skipped: the
    # oracle must agree with the FP64 restatement, which drops them explicitly
    q = np.random.default_rng(9).uniform(-2, 2, size=(50, dim)).astype(np.float32)
    np.testing.assert_allclose(O.min_distance(ctx, q), [O.min_distance64(cc, qi) for qi in q], atol=1e-5)


def test_disabled_context_is_collision_free():
    cc = _single_sphere_robot({})
    ctx = O.Context(cc, 1)
    assert not O.in_collision(ctx, np.zeros(1, np.float32))[0]


def _frame_groups(cc, n_frames):
    """Per frame: member sphere indices, local AABB-centre bounding sphere (radius + |r|)."""
    inv = np.asarray(cc["fk_topo_inv"])
    links = np.asarray(cc["sphere_link_idx"])
    loc = np.asarray(cc["sphere_local"], np.float32)
    rad = np.asarray(cc["sphere_radius"], np.float32)
    groups = []
    for j in range(n_frames):
        mem = [s for s, l in enumerate(links) if 0 <= l < n_frames and inv[l] == j]
        if not mem:
            continue
        p = loc[mem]
        c = (0.5 * (p.min(0) + p.max(0))).astype(np.float32)
        R = np.float32(np.max(np.linalg.norm(p - c, axis=1) + np.abs(rad[mem])))
        groups.append((j, mem, c, R))
    return groups


@pytest.mark.parametrize("name", ["c2", "tree"])
def test_coarse_culling_never_drops_a_hit(name):
    """The bound the culled GPU check relies on, in float32: whenever a frame's bounding sphere
    is farther from an obstacle than R + |r_o| + 1 mm, no sphere of that frame hits it (binary,
    squared form) nor comes within the margin."""
    cc, dim = _robots()[name]
    if name == "tree":
        cc = dict(cc)
        rng0 = np.random.default_rng(5)
        cc["world_spheres"] = np.concatenate([rng0.uniform(-0.5, 0.5, (40, 3)), rng0.uniform(0.01, 0.1, (40, 1))],
                                             1).astype(np.float32)
    ctx = O.Context(cc, dim)
    inv = np.asarray(cc["fk_topo_inv"])
    n = np.asarray(cc["fk_parent_idx"]).shape[0]
    groups = _frame_groups(cc, n)
    ws = np.asarray(cc["world_spheres"], np.float32).reshape(-1, 4)
    rad = np.asarray(cc["sphere_radius"], np.float32)
    pad, margin = np.float32(1e-3), np.float32(0.02)
    rng = np.random.default_rng(1)
    lo, hi = (-2.8, 2.8)
    culled = kept = 0
    for _ in range(120):
        q = rng.uniform(lo, hi, size=dim).astype(np.float32)
        T = O.fk(ctx, q)                     # [n_j, 7] FP32, original joint order
        P = O.sphere_positions(ctx, q)       # [n_rs, 3] FP32
        for j, mem, c, R in groups:
            fr = T[np.flatnonzero(inv == j)[0]]
            Rm = O.quat_mat64(fr[:4]).astype(np.float32)
            cw = (Rm @ c + fr[4:]).astype(np.float32)
            d = np.linalg.norm(ws[:, :3] - cw, axis=1).astype(np.float32)
            for mode_cut in (np.float32(0.0), margin):
                cull = d > R + np.abs(ws[:, 3]) + mode_cut + pad
                for o in np.flatnonzero(cull):
                    dd = np.linalg.norm(P[mem] - ws[o, :3], axis=1)
                    assert np.all(dd > np.abs(rad[mem] + ws[o, 3])), (j, o)          # no squared-form hit
                    assert np.all(dd - rad[mem] - ws[o, 3] >= mode_cut), (j, o)     # no distance < cut
                culled += int(cull.sum())
                kept += int((~cull).sum())
    assert culled > kept  # the bound prunes most pairs


@pytest.mark.parametrize("name", ["c2", "tree"])
def test_unit_culling_never_drops_a_hit(name):
    """The bound of the planner's sub-unit check, in float32: for up to 16 consecutive waypoints
    of an edge, a frame's unit sphere (midpoint of the box around the waypoints' bounding-sphere
    centres, radius R + cut + 1 mm + the box's half-diagonal x 1.000001) that is farther from an
    obstacle than that radius + |r_o| proves that no sphere of the frame hits the obstacle (binary,
    squared form) or comes within the margin at any of the waypoints.  Without the half-diagonal
    the same sample drops hits, so the spread term is what the bound rests on."""
    cc, dim = _robots()[name]
    cc = dict(cc)
    rng0 = np.random.default_rng(7)
    cc["world_spheres"] = np.concatenate([rng0.uniform(-0.6, 0.6, (60, 3)), rng0.uniform(0.01, 0.1, (60, 1))],
                                         1).astype(np.float32)
    ctx = O.Context(cc, dim)
    inv = np.asarray(cc["fk_topo_inv"])
    n = np.asarray(cc["fk_parent_idx"]).shape[0]
    groups = _frame_groups(cc, n)
    ws = np.asarray(cc["world_spheres"], np.float32).reshape(-1, 4)
    rad = np.asarray(cc["sphere_radius"], np.float32)
    pad, margin = np.float32(1e-3), np.float32(0.02)
    rng = np.random.default_rng(3)
    culled = kept = dropped_without_spread = 0
    for _ in range(60):
        a = rng.uniform(-2.8, 2.8, size=dim).astype(np.float32)
        d = rng.normal(size=dim).astype(np.float32)
        b = (a + np.float32(rng.uniform(0.1, 0.8)) * d / np.float32(np.linalg.norm(d))).astype(np.float32)
        nu = int(rng.integers(2, 17))
        qs = [(a + np.float32((k + 1) / 16.0) * (b - a)).astype(np.float32) for k in range(nu)]
        Ts = [O.fk(ctx, q) for q in qs]
        Ps = [O.sphere_positions(ctx, q) for q in qs]
        for j, mem, c, R in groups:
            row = np.flatnonzero(inv == j)[0]
            cws = np.stack([(O.quat_mat64(T[row, :4]).astype(np.float32) @ c + T[row, 4:]).astype(np.float32)
                            for T in Ts])
            lo, hi = cws.min(0), cws.max(0)
            cu = (np.float32(0.5) * (lo + hi)).astype(np.float32)
            e = (hi - lo).astype(np.float32)
            spread = np.float32(0.5) * np.float32(np.sqrt(np.float32(e @ e)))
            du = np.linalg.norm(ws[:, :3] - cu, axis=1).astype(np.float32)
            for cut in (np.float32(0.0), margin):
                lim = (R + cut + pad) + spread * np.float32(1.000001)
                cull = du > lim + np.abs(ws[:, 3])
                for o in np.flatnonzero(cull):
                    for P in Ps:
                        dd = np.linalg.norm(P[mem] - ws[o, :3], axis=1)
                        assert np.all(dd > np.abs(rad[mem] + ws[o, 3])), (j, o)
                        assert np.all(dd - rad[mem] - ws[o, 3] >= cut), (j, o)
                culled += int(cull.sum())
                kept += int((~cull).sum())
                # the bound without the spread: count the pairs it would clear wrongly
                bad = du > (R + cut + pad) + np.abs(ws[:, 3])
                for o in np.flatnonzero(bad & ~cull):
                    if any(np.any(np.linalg.norm(P[mem] - ws[o, :3], axis=1) - rad[mem] - ws[o, 3] < cut)
                           for P in Ps):
                        dropped_without_spread += 1
    assert culled > kept
    assert dropped_without_spread > 0
