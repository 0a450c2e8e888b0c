"""URDF/SRDF -> collision_context (the reference's utils.py:119-252 layout) on the CPU.

The synthetic arms of ``prrax_amd.scenes`` are written out as URDF text (spheres as collision
geometry, finger joints as ``<mimic>``), parsed back, and the built context must move every
sphere exactly as the original context does (FP64 FK of both), with the layout invariants of
the reference builder: spheres sorted by frame, coarse spheres enclosing their groups, self
pairs = full products of the active link pairs with contiguous ranges.  A shuffled file order
exercises the topological sort and ``fk_topo_inv``; the structural checks of
evaluate_mbm.py:70-224 reject each broken context they are meant to catch.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from prrax_amd import robot_model as RM
from prrax_amd import scenes, validate


def robot_urdf(robot: scenes.Robot, name: str = "arm", order=None, rpy=None) -> str:
    """URDF text of a synthetic robot (every joint's parent transform has identity rotation unless
    ``rpy`` gives per-joint roll-pitch-yaw, in which case the translation is kept and the rotation
    replaced)."""
    cc = robot.ctx
    n = cc["fk_parent_idx"].shape[0]
    act, mim = cc["fk_act_idx"], cc["fk_mimic_act_idx"]
    driver = {int(a): j for j, a in enumerate(act) if a >= 0}
    lines = [f'<robot name="{name}">', '<link name="base"/>']
    for j in range(n):
        sph = [s for s in range(cc["sphere_link_idx"].shape[0]) if cc["sphere_link_idx"][s] == j]
        lines.append(f'<link name="link{j}">')
        for s in sph:
            x, y, z = (float(v) for v in cc["sphere_local"][s])
            lines.append(f'  <collision><origin xyz="{x!r} {y!r} {z!r}"/><geometry>'
                         f'<sphere radius="{float(cc["sphere_radius"][s])!r}"/></geometry></collision>')
        lines.append("</link>")
    for j in (order if order is not None else range(n)):
        p = int(cc["fk_parent_idx"][j])
        t = cc["fk_parent_tf"][j]
        ax = cc["fk_twists"][j, 3:]
        r = rpy[j] if rpy is not None else (0.0, 0.0, 0.0)
        lines.append(f'<joint name="j{j}" type="revolute"><parent link="{"base" if p < 0 else f"link{p}"}"/>'
                     f'<child link="link{j}"/><origin xyz="{float(t[4])!r} {float(t[5])!r} {float(t[6])!r}" '
                     f'rpy="{r[0]!r} {r[1]!r} {r[2]!r}"/><axis xyz="{float(ax[0])!r} {float(ax[1])!r} '
                     f'{float(ax[2])!r}"/><limit lower="-2.8" upper="2.8"/>')
        if mim[j] >= 0:
            lines.append(f'  <mimic joint="j{driver[int(mim[j])]}" multiplier="{float(cc["fk_mimic_mul"][j])!r}" '
                         f'offset="{float(cc["fk_mimic_off"][j])!r}"/>')
        lines.append("</joint>")
    lines.append("</robot>")
    return "\n".join(lines)


def _positions(cc, q):
    P, r = validate.sphere_centres(cc, q)
    return P, r


@pytest.mark.parametrize("which", ["arm", "dual_arm"])
def test_urdf_round_trip_moves_every_sphere_the_same(which):
    robot = scenes.make_arm(0, 48, 60) if which == "arm" else scenes.make_dual_arm(0)
    cc, lo, hi, model = RM.load_robot(robot_urdf(robot))
    assert model.dim == robot.dim
    np.testing.assert_array_equal(lo, robot.lower)
    np.testing.assert_array_equal(hi, robot.upper)
    for k in ("fk_twists", "fk_parent_idx", "fk_act_idx", "fk_mimic_act_idx", "fk_mimic_mul", "fk_mimic_off"):
        np.testing.assert_array_equal(cc[k], robot.ctx[k], err_msg=k)
    np.testing.assert_allclose(cc["fk_parent_tf"], robot.ctx["fk_parent_tf"], atol=0)
    rng = np.random.default_rng(1)
    q = rng.uniform(robot.lower, robot.upper, size=(64, robot.dim))
    Pa, ra = _positions(cc, q)
    Pb, rb = _positions(robot.ctx, q)
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_allclose(Pa, Pb, rtol=0, atol=1e-12)
    # the FP32 oracle FK agrees with the original context bit for bit
    a = O.fk(O.Context(cc, robot.dim), q[0].astype(np.float32))
    b = O.fk(O.Context(robot.ctx, robot.dim), q[0].astype(np.float32))
    np.testing.assert_array_equal(a, b)


def test_shuffled_file_order_and_rotated_origins():
    """Joints listed children first and joint origins with roll-pitch-yaw: the sort restores the
    chain, fk_topo_inv maps file indices to it, and FK equals a direct FP64 evaluation of the URDF
    transforms."""
    robot = scenes.make_arm(2, 21, 10)
    n = 7
    order = [6, 2, 4, 0, 5, 1, 3]
    rpy = [(0.1 * j, -0.2 + 0.05 * j, 0.3 - 0.1 * j) for j in range(n)]
    cc, lo, hi, model = RM.load_robot(robot_urdf(robot, order=order, rpy=rpy))
    assert [j.name for j in model.joints] == [f"j{k}" for k in range(n)]  # topological = chain order
    # file index f holds joint order[f], which sits at position order[f] of the chain
    np.testing.assert_array_equal(cc["fk_topo_inv"], np.array(order, np.int32))
    # sphere_link_idx holds file indices: sphere on link j -> file index of joint j
    file_of = {j: order.index(j) for j in range(n)}
    want_links = sorted(file_of[int(j)] for j in robot.ctx["sphere_link_idx"])
    np.testing.assert_array_equal(cc["sphere_link_idx"], want_links)
    # direct FP64 FK of the URDF: T_j = T_parent . [R(rpy) | xyz] . Rot(axis, q)
    q = np.random.default_rng(3).uniform(-2.8, 2.8, size=7)
    T = np.eye(4)
    want = []
    for j in range(n):
        r, p, y = rpy[j]
        Rx = np.array([[1, 0, 0], [0, np.cos(r), -np.sin(r)], [0, np.sin(r), np.cos(r)]])
        Ry = np.array([[np.cos(p), 0, np.sin(p)], [0, 1, 0], [-np.sin(p), 0, np.cos(p)]])
        Rz = np.array([[np.cos(y), -np.sin(y), 0], [np.sin(y), np.cos(y), 0], [0, 0, 1]])
        P = np.eye(4)
        P[:3, :3] = Rz @ Ry @ Rx
        P[:3, 3] = robot.ctx["fk_parent_tf"][j, 4:]
        a = robot.ctx["fk_twists"][j, 3:].astype(np.float64)
        K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
        E = np.eye(4)
        E[:3, :3] = np.eye(3) + np.sin(q[j]) * K + (1 - np.cos(q[j])) * K @ K
        T = T @ P @ E
        want.append(T.copy())
    got = O.fk64(cc, q)  # [n_joints, 4, 4] in file order
    for j in range(n):
        # the context stores the origins' quaternions in float32 (1 ulp ~ 6e-8 per rotation)
        np.testing.assert_allclose(got[file_of[j]], want[j], atol=1e-6)


def test_builder_layout_invariants():
    robot = scenes.make_dual_arm(0)
    srdf = '<robot name="d"><disable_collisions link1="link0" link2="link2" reason="x"/>' \
           '<disable_collisions link1="link9" link2="link11" reason="x"/></robot>'
    cc, lo, hi, model = RM.load_robot(robot_urdf(robot), srdf)
    sl = cc["sphere_link_idx"]
    assert np.all(np.diff(sl) >= 0)  # sorted by frame
    off = cc["fine_link_offsets"]
    assert off[0] == 0 and off[-1] == sl.shape[0] and np.all(np.diff(off) > 0)
    for g in range(off.shape[0] - 1):
        members = slice(off[g], off[g + 1])
        assert np.all(sl[members] == cc["coarse_sphere_link_idx"][g])
        d = np.linalg.norm(cc["sphere_local"][members] - cc["coarse_sphere_local"][g], axis=1) + cc["sphere_radius"][members]
        assert np.all(d <= cc["coarse_sphere_radius"][g] + 1e-6)  # encloses every member
    sp, rng_, cps = cc["self_pairs"], cc["self_pair_ranges"], cc["coarse_self_pairs"]
    assert rng_[0, 0] == 0 and rng_[-1, 1] == sp.shape[0] and np.all(rng_[1:, 0] == rng_[:-1, 1])
    for (s, e), (ga, gb) in zip(rng_, cps):
        a, b = set(sp[s:e, 0].tolist()), set(sp[s:e, 1].tolist())
        assert len(a) * len(b) == e - s  # a full product
        assert all(off[ga] <= i < off[ga + 1] for i in a) and all(off[gb] <= i < off[gb + 1] for i in b)
    # disabled and adjacent link pairs have no block; every other pair of links with spheres does
    pairs = RM.active_link_pairs(model, RM.parse_srdf(srdf))
    assert ("link0", "link2") not in pairs and ("link0", "link1") not in pairs and ("link0", "link3") in pairs
    assert len(pairs) == cps.shape[0]


def test_verify_collision_context_catches_each_broken_context():
    robot = scenes.make_arm(0, 48, 60)
    urdf = robot_urdf(robot)
    # links 1 and 3 overlap when the arm is straight: without an SRDF entry every
    # configuration near zero would be in self-collision
    cc0, lo, hi, _ = RM.load_robot(urdf)
    with pytest.raises(RM.ContextError, match="zero configuration"):
        RM.verify_collision_context(cc0, lo, hi)
    srdf = '<robot name="arm"><disable_collisions link1="link1" link2="link3" reason="Never"/></robot>'
    cc, lo, hi, _ = RM.load_robot(urdf, srdf)
    RM.verify_collision_context(cc, lo, hi, start=np.zeros(7), goals=np.zeros((2, 7)))
    bad = [
        (dict(cc), lo, lo - 1.0, "limits"),
        (dict(cc, fk_act_idx=np.array([0, 1, 2, 3, 4, 5, 9], np.int32)), lo, hi, "fk_act_idx"),
        (dict(cc, fk_act_idx=np.array([0, 1, 2, 3, 4, 5, -1], np.int32)), lo, hi, "do not appear"),
        (dict(cc, fk_mimic_act_idx=np.array([-1, -1, -1, -1, -1, -1, 8], np.int32)), lo, hi, "mimic"),
        (dict(cc, sphere_link_idx=np.where(np.arange(48) == 3, 11, cc["sphere_link_idx"]).astype(np.int32)), lo, hi,
         "sphere_link_idx"),
    ]
    for c, l, h, msg in bad:
        with pytest.raises(RM.ContextError, match=msg):
            RM.verify_collision_context(c, l, h)
    # adjacent links are ignored by default; with them active, spheres touching across a joint
    # collide at q = 0 as well
    cc_adj, lo2, hi2, _ = RM.load_robot(urdf, srdf, ignore_adjacent=False)
    with pytest.raises(RM.ContextError, match="zero configuration"):
        RM.verify_collision_context(cc_adj, lo2, hi2)
    # a start folded onto itself is in self-collision
    folded = np.array([0, 2.6, 0, 2.6, 0, 2.6, 0.0])
    assert validate.self_pair_distances(cc, folded).min() <= -1e-3
    with pytest.raises(RM.ContextError, match="self-collision"):
        RM.verify_collision_context(cc, lo, hi, start=folded)
    with pytest.raises(RM.ContextError, match="dimension"):
        RM.verify_collision_context(cc, lo, hi, start=np.zeros(6))


def test_world_arrays_rows():
    w = RM.world_arrays([
        {"type": "sphere", "center": [1, 2, 3], "radius": 0.5},
        {"type": "capsule", "center": [0, 0, 1], "axis": [0, 0, 2], "height": 0.4, "radius": 0.1},
        {"type": "box", "center": [1, 0, 0], "half_lengths": [0.1, 0.2, 0.3], "quat": [np.cos(0.25), 0, 0, np.sin(0.25)]},
        {"type": "halfspace", "normal": [0, 0, 2], "point": [0, 0, -0.1]},
    ])
    np.testing.assert_allclose(w["world_spheres"], [[1, 2, 3, 0.5]])
    np.testing.assert_allclose(w["world_capsules"], [[0, 0, 0.8, 0, 0, 1.2, 0.1]], atol=1e-7)
    c, s = np.cos(0.5), np.sin(0.5)  # rotation by 0.5 rad about z: columns are the box axes
    np.testing.assert_allclose(w["world_boxes"], [[1, 0, 0, c, s, 0, -s, c, 0, 0, 0, 1, 0.1, 0.2, 0.3]], atol=1e-7)
    np.testing.assert_allclose(w["world_halfspaces"], [[0, 0, 1, 0, 0, -0.1]])
    with pytest.raises(ValueError, match="unknown obstacle"):
        RM.world_arrays([{"type": "mesh"}])


def test_suite_statistics():
    from prrax_amd.evaluate import SuiteResult, describe, markdown_table, summary

    d = describe([1.0, 2.0, 3.0, 4.0])
    assert d["mean"] == 2.5 and d["min"] == 1.0 and d["max"] == 4.0 and d["50%"] == 2.5
    assert abs(d["std"] - np.std([1, 2, 3, 4], ddof=1)) < 1e-12
    res = SuiteResult(rows=[{"set": "a", "planning_time_us": 10.0, "kernel_time_us": 5.0, "iterations": 1, "cost": 2.0},
                            {"set": "a", "planning_time_us": 30.0, "kernel_time_us": 7.0, "iterations": 3, "cost": 4.0}],
                      total=4, valid=3, failed=1)
    st = summary(res)
    assert (st["solved"], st["valid"], st["total"]) == (2, 3, 4)
    table = markdown_table(st, "all")
    assert "| mean | 20 | 6 | 2 | 3 |" in table
