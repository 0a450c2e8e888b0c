"""Robot descriptions (URDF + SRDF) to the planner's ``collision_context``.

The reference builds its context from pyroffi's robot model (``cuda-rrtc/jax/utils.py:119-252``,
``build_prrtc_collision_context``): twist FK arrays in topological joint order, spherized links,
SRDF-filtered self pairs, and the optional coarse-culling arrays.  pyroffi is not available
here, so this module reads the description files itself (standard-library XML only) and then
builds the same wire format:

* ``parse_urdf`` -- links, joints (revolute, continuous, prismatic, fixed; ``<mimic>``), joint
  origins (xyz + rpy, R = Rz(yaw) Ry(pitch) Rx(roll)), axes and limits, and each link's
  collision spheres (``<collision><geometry><sphere radius=...>``, centre = the collision
  origin).  Joints are sorted topologically (parents first, file order otherwise); the FK arrays
  are stored in that order and ``fk_topo_inv`` maps a joint's file index to it, as the reference
  kernel reads them (``sphere_link_idx`` holds file indices).
* ``parse_srdf`` -- the ``<disable_collisions link1= link2=>`` pairs.
* ``build_collision_context`` -- the layout of ``utils.py:119-252``: spheres with a positive
  radius, sorted (stably) by the FK frame they move with (the joint whose child is their link,
  -1 for the root link); one coarse sphere per frame group (centre of the members' bounding
  box, radius = the largest |c_i - c| + r_i) with ``fine_link_offsets``; self pairs = every fine
  pair of every active link pair (all pairs of links with spheres, minus the SRDF's and, by
  default, those joined by a joint), grouped per link pair with ``coarse_self_pairs`` and
  ``self_pair_ranges``; world primitives from ``world_arrays``.
* ``world_arrays`` -- obstacles (dicts) to the reference's primitive rows
  (``utils.py:58-116``): spheres ``[x, y, z, r]``, capsules ``[p1, p2, r]`` (centre +- axis
  height / 2), boxes ``[c, a1, a2, a3, half lengths]`` (a_i = rotation columns), halfspaces
  ``[n, p]``.
* ``verify_collision_context`` -- the structural checks of ``evaluate_mbm.py:70-224``.
"""

from __future__ import annotations

import math
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import Iterable, Optional, Sequence

import numpy as np

MAX_JOINTS = 64


@dataclass
class Joint:
    name: str
    type: str
    parent: str
    child: str
    xyz: np.ndarray
    rpy: np.ndarray
    axis: np.ndarray
    lower: float
    upper: float
    mimic: Optional[tuple] = None  # (joint name, multiplier, offset)


@dataclass
class RobotModel:
    name: str
    links: list                      # link names in file order
    joints: list                     # Joint, topological order
    file_index: list                 # file index of each sorted joint
    spheres: dict = field(default_factory=dict)  # link -> [(x, y, z, r), ...]

    @property
    def root(self) -> str:
        children = {j.child for j in self.joints}
        roots = [l for l in self.links if l not in children]
        return roots[0] if roots else self.links[0]

    @property
    def actuated(self) -> list:
        """Sorted indices of the actuated joints (moving, not mimic), in topological order."""
        return [k for k, j in enumerate(self.joints) if j.type != "fixed" and j.mimic is None]

    @property
    def dim(self) -> int:
        return len(self.actuated)

    def limits(self) -> tuple:
        act = self.actuated
        lo = np.array([self.joints[k].lower for k in act], np.float32)
        hi = np.array([self.joints[k].upper for k in act], np.float32)
        return lo, hi

    def parent_joint_of_link(self) -> dict:
        """Link -> file index of the joint whose child it is (-1 for the root)."""
        out = {l: -1 for l in self.links}
        for k, j in enumerate(self.joints):
            out[j.child] = self.file_index[k]
        return out

    def adjacent_links(self) -> set:
        return {frozenset((j.parent, j.child)) for j in self.joints}


def _floats(text: Optional[str], n: int, default: float = 0.0) -> np.ndarray:
    if text is None:
        return np.full(n, default)
    vals = [float(v) for v in text.split()]
    if len(vals) != n:
        raise ValueError(f"expected {n} numbers, got {text!r}")
    return np.array(vals)


def rpy_to_quat(rpy) -> np.ndarray:
    """(w, x, y, z) of R = Rz(yaw) Ry(pitch) Rx(roll) (URDF's fixed-axis roll-pitch-yaw)."""
    r, p, y = (0.5 * float(v) for v in rpy)
    cr, sr, cp, sp, cy, sy = math.cos(r), math.sin(r), math.cos(p), math.sin(p), math.cos(y), math.sin(y)
    return np.array([cr * cp * cy + sr * sp * sy, sr * cp * cy - cr * sp * sy, cr * sp * cy + sr * cp * sy,
                     cr * cp * sy - sr * sp * cy])


def parse_urdf(text: str) -> RobotModel:
    root = ET.fromstring(text)
    if root.tag != "robot":
        raise ValueError("not a URDF: the root element must be <robot>")
    links, spheres = [], {}
    for le in root.findall("link"):
        name = le.get("name")
        links.append(name)
        for ce in le.findall("collision"):
            ge = ce.find("geometry")
            se = ge.find("sphere") if ge is not None else None
            if se is None:
                continue
            oe = ce.find("origin")
            c = _floats(oe.get("xyz") if oe is not None else None, 3)
            spheres.setdefault(name, []).append((float(c[0]), float(c[1]), float(c[2]), float(se.get("radius"))))
    file_joints = []
    for je in root.findall("joint"):
        jt = je.get("type")
        if jt not in ("revolute", "continuous", "prismatic", "fixed"):
            raise ValueError(f"joint {je.get('name')!r}: type {jt!r} is not supported")
        oe, ae, lim, me = je.find("origin"), je.find("axis"), je.find("limit"), je.find("mimic")
        axis = _floats(ae.get("xyz") if ae is not None else None, 3, 0.0) if ae is not None else np.array([1.0, 0, 0])
        if jt != "fixed":
            n = np.linalg.norm(axis)
            if n == 0.0:
                raise ValueError(f"joint {je.get('name')!r}: zero axis")
            axis = axis / n
        if jt == "continuous":
            lo, hi = -math.pi, math.pi
        elif lim is not None:
            lo, hi = float(lim.get("lower", 0.0)), float(lim.get("upper", 0.0))
        else:
            lo, hi = 0.0, 0.0
        mimic = None
        if me is not None:
            mimic = (me.get("joint"), float(me.get("multiplier", 1.0)), float(me.get("offset", 0.0)))
        file_joints.append(Joint(je.get("name"), jt, je.find("parent").get("link"), je.find("child").get("link"),
                                 _floats(oe.get("xyz") if oe is not None else None, 3),
                                 _floats(oe.get("rpy") if oe is not None else None, 3), axis, lo, hi, mimic))
    known = set(links)
    for j in file_joints:
        if j.parent not in known or j.child not in known:
            raise ValueError(f"joint {j.name!r} refers to an unknown link")
    # topological order: a joint after the joint that produces its parent link; file order otherwise
    producer = {j.child: k for k, j in enumerate(file_joints)}
    if len(producer) != len(file_joints):
        raise ValueError("a link is the child of two joints")
    order, placed = [], set()
    while len(order) < len(file_joints):
        progressed = False
        for k, j in enumerate(file_joints):
            if k in placed:
                continue
            p = producer.get(j.parent)
            if p is None or p in placed:
                order.append(k)
                placed.add(k)
                progressed = True
        if not progressed:
            raise ValueError("the joints form a cycle")
    return RobotModel(root.get("name", ""), links, [file_joints[k] for k in order], order, spheres)


def parse_srdf(text: str) -> set:
    """Link pairs whose collisions the SRDF disables (``<disable_collisions>``)."""
    root = ET.fromstring(text)
    return {frozenset((e.get("link1"), e.get("link2"))) for e in root.iter("disable_collisions")}


def fk_arrays(model: RobotModel) -> dict:
    """The context's FK keys (sorted joint order) for a parsed robot."""
    n = len(model.joints)
    if n > MAX_JOINTS:
        raise ValueError(f"{n} joints exceed the planner's cap of {MAX_JOINTS}")
    sorted_of_file = {f: k for k, f in enumerate(model.file_index)}
    producer = {j.child: k for k, j in enumerate(model.joints)}
    act_of = {k: a for a, k in enumerate(model.actuated)}
    by_name = {j.name: k for k, j in enumerate(model.joints)}
    tw = np.zeros((n, 6), np.float32)
    ptf = np.zeros((n, 7), np.float32)
    par = np.full(n, -1, np.int32)
    act = np.full(n, -1, np.int32)
    mul = np.ones(n, np.float32)
    off = np.zeros(n, np.float32)
    mim = np.full(n, -1, np.int32)
    for k, j in enumerate(model.joints):
        if j.type in ("revolute", "continuous"):
            tw[k, 3:] = j.axis
        elif j.type == "prismatic":
            tw[k, :3] = j.axis
        ptf[k, :4] = rpy_to_quat(j.rpy)
        ptf[k, 4:] = j.xyz
        par[k] = producer.get(j.parent, -1)
        if j.mimic is not None:
            # follow a mimic chain to the actuated joint it ends at, composing the affine maps
            m, o, src = j.mimic[1], j.mimic[2], by_name.get(j.mimic[0])
            seen = {k}
            while src is not None and model.joints[src].mimic is not None and src not in seen:
                seen.add(src)
                nm, no, nxt = model.joints[src].mimic[1], model.joints[src].mimic[2], by_name.get(model.joints[src].mimic[0])
                m, o, src = m * nm, m * no + o, nxt
            if src is None or src not in act_of:
                raise ValueError(f"joint {j.name!r} mimics a joint that is not actuated")
            mim[k], mul[k], off[k] = act_of[src], m, o
        elif j.type != "fixed":
            act[k] = act_of[k]
    inv = np.array([sorted_of_file[f] for f in range(n)], np.int32)
    return {"fk_twists": tw, "fk_parent_tf": ptf, "fk_parent_idx": par, "fk_act_idx": act, "fk_mimic_mul": mul,
            "fk_mimic_off": off, "fk_mimic_act_idx": mim, "fk_topo_inv": inv}


def active_link_pairs(model: RobotModel, disabled: Iterable = (), ignore_adjacent: bool = True) -> list:
    """(link i, link j), i before j in file order: links with spheres, not disabled, not adjacent."""
    dis = {frozenset(p) for p in disabled}
    adj = model.adjacent_links() if ignore_adjacent else set()
    with_spheres = [l for l in model.links if model.spheres.get(l)]
    out = []
    for a in range(len(with_spheres)):
        for b in range(a + 1, len(with_spheres)):
            pair = frozenset((with_spheres[a], with_spheres[b]))
            if pair in dis or pair in adj:
                continue
            out.append((with_spheres[a], with_spheres[b]))
    return out


def _vec(o: dict, key: str, n: int) -> np.ndarray:
    v = np.asarray(o[key], np.float64).reshape(-1)
    if v.shape[0] != n:
        raise ValueError(f"obstacle {key!r} needs {n} values")
    return v


def world_arrays(obstacles: Sequence[dict] = ()) -> dict:
    """Obstacle dicts -> the context's world keys.  Types: ``sphere`` (center, radius),
    ``capsule`` (center, axis, height, radius), ``box`` (center, half_lengths, rotation 3x3 or
    quat (w, x, y, z); identity when absent), ``halfspace`` (normal, point; free side +normal)."""
    sph, cap, box, hs = [], [], [], []
    for o in obstacles:
        t = o.get("type")
        if t == "sphere":
            sph.append([*_vec(o, "center", 3), float(o["radius"])])
        elif t == "capsule":
            c, ax = _vec(o, "center", 3), _vec(o, "axis", 3)
            ax = ax / np.linalg.norm(ax)
            h = 0.5 * float(o["height"])
            cap.append([*(c - h * ax), *(c + h * ax), float(o["radius"])])
        elif t == "box":
            c, hl = _vec(o, "center", 3), _vec(o, "half_lengths", 3)
            if "rotation" in o:
                R = np.asarray(o["rotation"], np.float64).reshape(3, 3)
            elif "quat" in o:
                w, x, y, z = _vec(o, "quat", 4) / np.linalg.norm(_vec(o, "quat", 4))
                R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                              [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                              [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
            else:
                R = np.eye(3)
            box.append([*c, *R[:, 0], *R[:, 1], *R[:, 2], *hl])
        elif t == "halfspace":
            n = _vec(o, "normal", 3)
            hs.append([*(n / np.linalg.norm(n)), *_vec(o, "point", 3)])
        else:
            raise ValueError(f"unknown obstacle type {t!r}")

    def rows(v, f):
        return np.asarray(v, np.float32).reshape(-1, f) if v else np.zeros((0, f), np.float32)

    return {"world_spheres": rows(sph, 4), "world_capsules": rows(cap, 7), "world_boxes": rows(box, 15),
            "world_halfspaces": rows(hs, 6)}


def build_collision_context(model: RobotModel, disabled: Iterable = (), obstacles: Sequence[dict] = (),
                            ignore_adjacent: bool = True) -> dict:
    """The full ``collision_context`` (utils.py:119-252 layout) of a parsed robot + obstacles."""
    ctx = fk_arrays(model)
    pj = model.parent_joint_of_link()
    # fine spheres in (link file order, sphere order), radius > 0
    raw = [(l, s) for l in model.links for s in model.spheres.get(l, []) if s[3] > 0.0]
    joint_per = np.array([pj[l] for l, _ in raw], np.int32)
    order = np.argsort(joint_per, kind="stable")
    ctx["sphere_link_idx"] = joint_per[order].astype(np.int32) if raw else np.zeros(0, np.int32)
    ctx["sphere_local"] = (np.array([s[:3] for _, s in raw], np.float32)[order] if raw
                           else np.zeros((0, 3), np.float32))
    ctx["sphere_radius"] = np.array([s[3] for _, s in raw], np.float32)[order] if raw else np.zeros(0, np.float32)
    new_of_old = np.empty(len(raw), np.int64)
    new_of_old[order] = np.arange(len(raw))
    # contiguous frame groups -> one coarse bounding sphere each
    sl = ctx["sphere_link_idx"]
    if sl.shape[0]:
        b = np.flatnonzero(np.diff(sl)) + 1
        starts, ends = np.concatenate(([0], b)), np.concatenate((b, [sl.shape[0]]))
    else:
        starts = ends = np.zeros(0, np.int64)
    ctx["coarse_sphere_link_idx"] = sl[starts].astype(np.int32) if sl.shape[0] else np.zeros(0, np.int32)
    ctx["fine_link_offsets"] = np.concatenate((starts, [sl.shape[0]])).astype(np.int32)
    cl = np.zeros((starts.shape[0], 3), np.float32)
    cr = np.zeros(starts.shape[0], np.float32)
    for k, (s, e) in enumerate(zip(starts, ends)):
        loc = ctx["sphere_local"][s:e].astype(np.float64)
        c = 0.5 * (loc.min(axis=0) + loc.max(axis=0))
        cl[k] = c
        cr[k] = float(np.max(np.linalg.norm(loc - c, axis=1) + ctx["sphere_radius"][s:e]))
    ctx["coarse_sphere_local"], ctx["coarse_sphere_radius"] = cl, cr
    group_of_joint = {int(j): k for k, j in enumerate(ctx["coarse_sphere_link_idx"])}
    # self pairs: every fine pair of each active link pair, one [start, end) block per link pair
    idx_of_link = {}
    for old, (l, _) in enumerate(raw):
        idx_of_link.setdefault(l, []).append(int(new_of_old[old]))
    blocks, coarse_pairs, ranges, cursor = [], [], [], 0
    for la, lb in active_link_pairs(model, disabled, ignore_adjacent):
        a, b = idx_of_link.get(la, []), idx_of_link.get(lb, [])
        if not a or not b:
            continue
        ga, gb = group_of_joint.get(pj[la], -1), group_of_joint.get(pj[lb], -1)
        if ga < 0 or gb < 0:
            continue
        blk = np.stack(np.meshgrid(np.array(a), np.array(b), indexing="ij"), axis=-1).reshape(-1, 2)
        blocks.append(blk)
        coarse_pairs.append((ga, gb))
        ranges.append((cursor, cursor + blk.shape[0]))
        cursor += blk.shape[0]
    ctx["self_pairs"] = np.concatenate(blocks).astype(np.int32) if blocks else np.zeros((0, 2), np.int32)
    ctx["coarse_self_pairs"] = np.asarray(coarse_pairs, np.int32).reshape(-1, 2)
    ctx["self_pair_ranges"] = np.asarray(ranges, np.int32).reshape(-1, 2)
    ctx.update(world_arrays(obstacles))
    return ctx


def load_robot(urdf_text: str, srdf_text: Optional[str] = None, obstacles: Sequence[dict] = (),
               ignore_adjacent: bool = True) -> tuple:
    """(collision_context, lower limits, upper limits, model) of a URDF (+ SRDF) description."""
    model = parse_urdf(urdf_text)
    disabled = parse_srdf(srdf_text) if srdf_text else set()
    cc = build_collision_context(model, disabled, obstacles, ignore_adjacent)
    lo, hi = model.limits()
    return cc, lo, hi, model


class ContextError(ValueError):
    """A collision context that would make the planner misbehave (evaluate_mbm.py:70-224)."""


def verify_collision_context(cc: dict, lower, upper, *, start=None, goals=None, name: str = "robot") -> None:
    """The checks of ``evaluate_mbm.py:70-224`` (raising ``ContextError``):

    1. lower < upper; 2. the limits, start and goals have n_act = len(lower) entries; 3.
    fk_act_idx in {-1} u [0, n_act); 4. every actuated index appears in fk_act_idx; 5.
    fk_mimic_act_idx in {-1} u [0, n_act); 6. sphere_link_idx in [-1, n_joints); 7. n_joints <=
    64; 8. no self pair in contact at the zero configuration; 9. start and goals free of
    self-collision (smallest self distance > -1e-3).  Distances are the host FP64 checker's.
    """
    from .validate import self_pair_distances

    lo, hi = np.asarray(lower, np.float64).reshape(-1), np.asarray(upper, np.float64).reshape(-1)
    act = np.asarray(cc["fk_act_idx"]).reshape(-1)
    mim = np.asarray(cc["fk_mimic_act_idx"]).reshape(-1)
    links = np.asarray(cc["sphere_link_idx"]).reshape(-1)
    nj = np.asarray(cc["fk_parent_idx"]).reshape(-1).shape[0]
    n_act = lo.shape[0]
    if not np.all(lo < hi):
        raise ContextError(f"[{name}] joint limits violated: lower >= upper for joints {np.where(lo >= hi)[0].tolist()}")
    if hi.shape[0] != n_act:
        raise ContextError(f"[{name}] upper has {hi.shape[0]} entries, lower {n_act}")
    for label, q in (("start", start), ("goals", goals)):
        if q is not None and np.asarray(q).shape[-1] != n_act:
            raise ContextError(f"[{name}] {label} dimension {np.asarray(q).shape[-1]} != n_act {n_act}")
    if not np.all((act == -1) | ((act >= 0) & (act < n_act))):
        raise ContextError(f"[{name}] fk_act_idx contains values outside [-1, n_act={n_act})")
    missing = sorted(set(range(n_act)) - set(act[act >= 0].tolist()))
    if missing:
        raise ContextError(f"[{name}] actuated indices {missing} do not appear in fk_act_idx")
    if not np.all((mim == -1) | ((mim >= 0) & (mim < n_act))):
        raise ContextError(f"[{name}] fk_mimic_act_idx contains values outside [-1, n_act={n_act})")
    if links.size and not np.all((links >= -1) & (links < nj)):
        raise ContextError(f"[{name}] sphere_link_idx contains values outside [-1, n_joints={nj})")
    if nj > MAX_JOINTS:
        raise ContextError(f"[{name}] n_joints={nj} exceeds the planner's cap of {MAX_JOINTS}")
    d0 = self_pair_distances(cc, np.zeros(n_act))
    bad = np.flatnonzero(d0 < 0.0)
    if bad.size:
        sp = np.asarray(cc["self_pairs"]).reshape(-1, 2)
        raise ContextError(f"[{name}] {bad.size} self pair(s) in contact at the zero configuration (e.g. spheres "
                           f"{sp[bad[0]].tolist()}, distance {d0[bad[0]]:.6f}): add SRDF disable_collisions entries")
    for label, q in (("start", start), ("goal", goals)):
        if q is None:
            continue
        for k, qq in enumerate(np.asarray(q, np.float64).reshape(-1, n_act)):
            d = self_pair_distances(cc, qq)
            if d.size and d.min() <= -1e-3:
                raise ContextError(f"[{name}] {label}[{k}] is in self-collision (smallest distance {d.min():.6f})")
