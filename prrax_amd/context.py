"""``collision_context`` checking, upload and batching.

The key set and row layouts are the reference's wire format (README.md:77-104,
cuda-rrtc/jax/prrtc.py:397-459, cuda-rrtc/jax/utils.py:119-252).  ``check_context`` applies the
structural invariants of ``evaluate_mbm.py:102-171`` (index ranges, joint cap) plus the ones the
FK needs to be well defined (parents before children, ``topo_inv`` a permutation); the
reference's kernel would instead silently treat ``n_joints > 64`` as collision-free
(prrtc_planner.cu:506).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib

REQUIRED_KEYS = (
    "fk_twists",
    "fk_parent_tf",
    "fk_parent_idx",
    "fk_act_idx",
    "fk_mimic_mul",
    "fk_mimic_off",
    "fk_mimic_act_idx",
    "fk_topo_inv",
    "sphere_link_idx",
    "sphere_local",
    "sphere_radius",
)
WORLD_SPECS = (("world_spheres", 4), ("world_capsules", 7), ("world_boxes", 15), ("world_halfspaces", 6))
# The optional coarse-checker keys (prrtc.py:76-110): key, row width (0 = a vector), dtype.
COARSE_SPECS = (("coarse_sphere_link_idx", 0, np.int32), ("coarse_sphere_local", 3, np.float32),
                ("coarse_sphere_radius", 0, np.float32), ("fine_link_offsets", 0, np.int32),
                ("coarse_self_pairs", 2, np.int32), ("self_pair_ranges", 2, np.int32))
MAX_JOINTS = 64
MAX_DIM = 16


def as_numpy(x, dtype) -> np.ndarray:
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(x, dtype=dtype))


def _rows(cc: dict, key: str, feat: int, dtype) -> np.ndarray:
    if key not in cc or cc[key] is None:
        return np.zeros((0, feat), dtype)
    a = as_numpy(cc[key], dtype)
    if a.size == 0:
        return np.zeros((0, feat), dtype)
    if a.ndim != 2 or a.shape[1] != feat:
        raise ValueError(f"collision_context['{key}'] must have shape (N, {feat}), got {a.shape}")
    return a


def check_context(cc: dict, dim: int) -> dict:
    """Validates a collision_context and returns contiguous numpy copies of its arrays."""
    missing = [k for k in REQUIRED_KEYS if k not in cc]
    if missing:
        raise ValueError(f"collision_context missing required keys: {missing}")
    out = {}
    out["fk_twists"] = _rows(cc, "fk_twists", 6, np.float32)
    out["fk_parent_tf"] = _rows(cc, "fk_parent_tf", 7, np.float32)
    nj = out["fk_twists"].shape[0]
    for k, dt in (("fk_parent_idx", np.int32), ("fk_act_idx", np.int32), ("fk_mimic_mul", np.float32),
                  ("fk_mimic_off", np.float32), ("fk_mimic_act_idx", np.int32), ("fk_topo_inv", np.int32)):
        a = as_numpy(cc[k], dt).reshape(-1)
        if a.shape[0] != nj:
            raise ValueError(f"collision_context['{k}'] has {a.shape[0]} entries, expected n_joints={nj}")
        out[k] = a
    if out["fk_parent_tf"].shape[0] != nj:
        raise ValueError("fk_parent_tf and fk_twists disagree on n_joints")
    if nj > MAX_JOINTS:
        raise ValueError(f"n_joints={nj} exceeds the planner's joint cap {MAX_JOINTS}")
    par = out["fk_parent_idx"]
    if np.any((par < -1) | (par >= np.arange(nj))):
        raise ValueError("fk_parent_idx must list each joint's parent before the joint (-1 for a root)")
    for k in ("fk_act_idx", "fk_mimic_act_idx"):
        a = out[k]
        if np.any((a != -1) & ((a < 0) | (a >= dim))):
            raise ValueError(f"collision_context['{k}'] contains values outside [-1, dim={dim})")
    if nj and not np.array_equal(np.sort(out["fk_topo_inv"]), np.arange(nj)):
        raise ValueError("fk_topo_inv must be a permutation of 0..n_joints-1")
    links = as_numpy(cc["sphere_link_idx"], np.int32).reshape(-1)
    nrs = links.shape[0]
    loc = as_numpy(cc["sphere_local"], np.float32).reshape(-1, 3) if nrs else np.zeros((0, 3), np.float32)
    rad = as_numpy(cc["sphere_radius"], np.float32).reshape(-1)
    if loc.shape[0] != nrs or rad.shape[0] != nrs:
        raise ValueError("sphere_link_idx, sphere_local and sphere_radius disagree on the sphere count")
    if np.any((links < -1) | (links >= max(nj, 0))) and nrs:
        raise ValueError(f"sphere_link_idx contains values outside [-1, n_joints={nj})")
    out["sphere_link_idx"], out["sphere_local"], out["sphere_radius"] = links, loc, rad
    for k, f in WORLD_SPECS:
        out[k] = _rows(cc, k, f, np.float32)
    out["self_pairs"] = _rows(cc, "self_pairs", 2, np.int32)
    for k, f, dt in COARSE_SPECS:
        if f:
            out[k] = _rows(cc, k, f, dt)
        else:
            out[k] = as_numpy(cc[k], dt).reshape(-1) if k in cc and cc[k] is not None else np.zeros(0, dt)
    nc = out["coarse_sphere_link_idx"].shape[0]
    if out["coarse_sphere_local"].shape[0] != nc or out["coarse_sphere_radius"].shape[0] != nc:
        raise ValueError("coarse_sphere_link_idx, coarse_sphere_local and coarse_sphere_radius disagree on the "
                         "coarse sphere count")
    return out


@dataclass
class DeviceRobot:
    tensors: dict
    struct: _lib.Robot
    n_joints: int
    n_spheres: int


@dataclass
class DeviceWorld:
    tensors: dict
    struct: _lib.World
    per_query: bool = False
    counts: np.ndarray | None = None
    keep: list = field(default_factory=list)


def _ptr(t: torch.Tensor | None) -> int | None:
    return t.data_ptr() if (t is not None and t.numel() > 0) else None


def upload_robot(arrays: dict, device) -> DeviceRobot:
    t = {k: torch.from_numpy(np.ascontiguousarray(arrays[k])).to(device) for k in REQUIRED_KEYS}
    s = _lib.Robot(
        _ptr(t["fk_twists"]), _ptr(t["fk_parent_tf"]), _ptr(t["fk_parent_idx"]), _ptr(t["fk_act_idx"]),
        _ptr(t["fk_mimic_mul"]), _ptr(t["fk_mimic_off"]), _ptr(t["fk_mimic_act_idx"]), _ptr(t["fk_topo_inv"]),
        int(arrays["fk_parent_idx"].shape[0]),
        _ptr(t["sphere_link_idx"]), _ptr(t["sphere_local"]), _ptr(t["sphere_radius"]),
        int(arrays["sphere_link_idx"].shape[0]),
    )
    return DeviceRobot(t, s, int(arrays["fk_parent_idx"].shape[0]), int(arrays["sphere_link_idx"].shape[0]))


COUNT_KEYS = ("world_spheres_count", "world_capsules_count", "world_boxes_count", "world_halfspaces_count",
              "self_pairs_count")


def _coarse_fields(arrays: dict, device, t: dict) -> list:
    """Device copies of the six coarse arrays (kept in ``t``) as the struct's pointer and count fields."""
    for k, _, _ in COARSE_SPECS:
        t[k] = torch.from_numpy(np.ascontiguousarray(arrays[k])).to(device)
    return [*(_ptr(t[k]) for k, _, _ in COARSE_SPECS), int(arrays["coarse_sphere_link_idx"].shape[0]),
            int(arrays["fine_link_offsets"].shape[0]), int(arrays["coarse_self_pairs"].shape[0]),
            int(arrays["self_pair_ranges"].shape[0])]


def upload_world(arrays: dict, device) -> DeviceWorld:
    keys = [k for k, _ in WORLD_SPECS] + ["self_pairs"]
    t = {k: torch.from_numpy(np.ascontiguousarray(arrays[k])).to(device) for k in keys}
    s = _lib.World(
        _ptr(t["world_spheres"]), _ptr(t["world_capsules"]), _ptr(t["world_boxes"]), _ptr(t["world_halfspaces"]),
        _ptr(t["self_pairs"]),
        *(int(arrays[k].shape[0]) for k in keys),
        None, None, None, None, None,
        *_coarse_fields(arrays, device, t),
    )
    return DeviceWorld(t, s)


def pack_worlds(worlds: list[dict], device, coarse: dict | None = None) -> DeviceWorld:
    """Per-query worlds padded to ``[B, max, F]`` plus the five per-query count vectors ``[B]``
    (the reference's batch-ctx arguments, prrtc_planner.cu:1964-1974; padding at
    prrtc.py:1009-1067).  ``coarse``: the checked arrays of the context the robot comes from."""
    B = len(worlds)
    keys = [(k, f, np.float32) for k, f in WORLD_SPECS] + [("self_pairs", 2, np.int32)]
    counts = np.zeros((B, 5), np.int32)
    tens = {}
    caps = []
    for col, (k, f, dt) in enumerate(keys):
        rows = [w[k] for w in worlds]
        cap = max((r.shape[0] for r in rows), default=0)
        packed = np.zeros((B, max(cap, 0), f), dt)
        for i, r in enumerate(rows):
            counts[i, col] = r.shape[0]
            if r.shape[0]:
                packed[i, : r.shape[0]] = r
        tens[k] = torch.from_numpy(packed).to(device)
        caps.append(cap)
    for col, k in enumerate(COUNT_KEYS):
        tens[k] = torch.from_numpy(np.ascontiguousarray(counts[:, col])).to(device)
    s = _lib.World(
        _ptr(tens["world_spheres"]), _ptr(tens["world_capsules"]), _ptr(tens["world_boxes"]),
        _ptr(tens["world_halfspaces"]), _ptr(tens["self_pairs"]), *caps,
        *(tens[k].data_ptr() for k in COUNT_KEYS),
        *_coarse_fields(coarse if coarse is not None else worlds[0], device, tens),
    )
    return DeviceWorld(tens, s, per_query=True, counts=counts)


def ctypes_ref(s: ctypes.Structure):
    return ctypes.byref(s)
