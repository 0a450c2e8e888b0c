"""ctypes binding of ``_prrtc_planner_lib.so`` (declarations in ``include/prrtc_amd.h``).

The library is the only compute path: if it is missing, fails to load or was built from other
sources than the ones next to it (``prrtc_version()`` carries the build's source hash), every
entry point raises ``RuntimeError`` (there is no CPU fallback).  Loading mirrors the reference's
``_load_and_register`` (cuda-rrtc/jax/prrtc.py:113-146): one ``ctypes.CDLL`` of the library
next to the package, cached for the process.
"""

from __future__ import annotations

import ctypes
from functools import lru_cache
from pathlib import Path

LIB_NAME = "_prrtc_planner_lib.so"
LIB_PATH = Path(__file__).resolve().parent / LIB_NAME

_f32p = ctypes.c_void_p
_vp = ctypes.c_void_p

PRRTC_OK = 0
PRRTC_E_INVALID = -1
PRRTC_E_HIP = -2
PRRTC_E_WORKSPACE = -3
PRRTC_E_LDS = -4

STATUS_SOLVED = 1
STATUS_ITER_CAP = 2
STATUS_TREE_FULL = 3

EXPORTED_SYMBOLS = (
    "prrtc_planner_workspace_bytes",
    "prrtc_planner_stats_offset",
    "prrtc_planner",
    "prrtc_planner_noalloc",
    "prrtc_planner_batch",
    "prrtc_check_configs",
    "prrtc_forward_kinematics",
    "prrtc_nearest_neighbor",
    "prrtc_extend",
    "prrtc_sphere_cost",
    "prrtc_refine_workspace_bytes",
    "prrtc_refine_paths",
    "prrtc_cloud_bytes",
    "prrtc_cloud_build",
    "prrtc_check_edges",
    "prrtc_last_error",
    "prrtc_version",
)


class Robot(ctypes.Structure):
    _fields_ = [
        ("fk_twists", _vp), ("fk_parent_tf", _vp), ("fk_parent_idx", _vp), ("fk_act_idx", _vp),
        ("fk_mimic_mul", _vp), ("fk_mimic_off", _vp), ("fk_mimic_act_idx", _vp), ("fk_topo_inv", _vp),
        ("n_joints", ctypes.c_int32),
        ("sphere_link_idx", _vp), ("sphere_local", _vp), ("sphere_radius", _vp),
        ("n_robot_spheres", ctypes.c_int32),
    ]


class World(ctypes.Structure):
    _fields_ = [
        ("world_spheres", _vp), ("world_capsules", _vp), ("world_boxes", _vp), ("world_halfspaces", _vp),
        ("self_pairs", _vp),
        ("n_world_spheres", ctypes.c_int32), ("n_world_capsules", ctypes.c_int32),
        ("n_world_boxes", ctypes.c_int32), ("n_world_halfspaces", ctypes.c_int32),
        ("n_self_pairs", ctypes.c_int32),
        ("world_spheres_count", _vp), ("world_capsules_count", _vp), ("world_boxes_count", _vp),
        ("world_halfspaces_count", _vp), ("self_pairs_count", _vp),
        ("coarse_sphere_link_idx", _vp), ("coarse_sphere_local", _vp), ("coarse_sphere_radius", _vp),
        ("fine_link_offsets", _vp), ("coarse_self_pairs", _vp), ("self_pair_ranges", _vp),
        ("n_coarse_spheres", ctypes.c_int32), ("n_fine_link_offsets", ctypes.c_int32),
        ("n_coarse_self_pairs", ctypes.c_int32), ("n_self_pair_ranges", ctypes.c_int32),
    ]


class Cloud(ctypes.Structure):
    _fields_ = [
        ("points", _vp), ("cell_start", _vp), ("n_points", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("lo", ctypes.c_float * 3), ("cell", ctypes.c_float), ("r_env", ctypes.c_float),
    ]


class Attrs(ctypes.Structure):
    _fields_ = [
        ("max_iterations", ctypes.c_int32), ("step_size", ctypes.c_float), ("num_new_samples", ctypes.c_int32),
        ("balance_mode", ctypes.c_int32), ("tree_ratio", ctypes.c_float), ("dynamic_domain", ctypes.c_int32),
        ("dd_alpha", ctypes.c_float), ("dd_radius", ctypes.c_float), ("dd_min_radius", ctypes.c_float),
        ("dim", ctypes.c_int32), ("max_nodes", ctypes.c_int32), ("granularity", ctypes.c_int32),
        ("collision_mode", ctypes.c_int32), ("collision_margin", ctypes.c_float),
    ]


class Outputs(ctypes.Structure):
    _fields_ = [
        ("tree_a_configs", _vp), ("tree_b_configs", _vp), ("tree_a_parents", _vp), ("tree_b_parents", _vp),
        ("tree_sizes", _vp), ("completed", _vp), ("iter_count", _vp), ("connection_info", _vp),
        ("solved_flag", _vp), ("kernel_time_ms", _vp),
        ("status", _vp), ("path", _vp), ("path_len", _vp), ("path_capacity", ctypes.c_int32),
    ]


class Launch(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_int32), ("slots", ctypes.c_int32), ("pass1_rounds", ctypes.c_int32),
                ("team_max", ctypes.c_int32), ("epoch_rounds", ctypes.c_int32), ("fk_batch", ctypes.c_int32)]


class PrrtcError(RuntimeError):
    """A library call returned an error code."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


@lru_cache(maxsize=1)
def lib() -> ctypes.CDLL:
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"pRRTC library not found at {LIB_PATH}. Build it with `python -m prrax_amd.build` "
            "(hipcc --offload-arch=gfx950)."
        )
    L = ctypes.CDLL(str(LIB_PATH))
    i32, sz, P = ctypes.c_int32, ctypes.c_size_t, ctypes.POINTER
    L.prrtc_version.argtypes = []
    L.prrtc_version.restype = ctypes.c_char_p
    _check_fresh(L.prrtc_version().decode())
    L.prrtc_planner_workspace_bytes.argtypes = [i32, i32, i32, i32]
    L.prrtc_planner_workspace_bytes.restype = sz
    L.prrtc_planner_stats_offset.argtypes = [i32, i32, i32, i32]
    L.prrtc_planner_stats_offset.restype = sz
    L.prrtc_planner.argtypes = [_vp, _vp, _vp, i32, _vp, _vp, P(Robot), P(World), P(Attrs), P(Outputs), P(Launch),
                                _vp, sz]
    L.prrtc_planner.restype = ctypes.c_int
    L.prrtc_planner_noalloc.argtypes = [_vp, _vp, _vp, i32, _vp, _vp, P(Robot), P(World), P(Attrs), P(Outputs),
                                        P(Launch)]
    L.prrtc_planner_noalloc.restype = ctypes.c_int
    L.prrtc_planner_batch.argtypes = [_vp, i32, _vp, _vp, i32, _vp, _vp, _vp, P(Robot), P(World), i32, _vp, _vp,
                                      _vp, P(Attrs), P(Outputs), P(Launch), _vp, sz]
    L.prrtc_planner_batch.restype = ctypes.c_int
    L.prrtc_check_configs.argtypes = [_vp, _vp, i32, i32, P(Robot), P(World), i32, ctypes.c_float, _vp, _vp,
                                      P(Launch), _vp]
    L.prrtc_check_configs.restype = ctypes.c_int
    L.prrtc_forward_kinematics.argtypes = [_vp, _vp, i32, i32, P(Robot), _vp]
    L.prrtc_forward_kinematics.restype = ctypes.c_int
    L.prrtc_nearest_neighbor.argtypes = [_vp, _vp, i32, i32, i32, _vp, i32, _vp, _vp]
    L.prrtc_nearest_neighbor.restype = ctypes.c_int
    L.prrtc_extend.argtypes = [_vp, _vp, i32, i32, _vp, _vp, i32, ctypes.c_float, _vp, _vp, _vp]
    L.prrtc_extend.restype = ctypes.c_int
    f32 = ctypes.c_float
    L.prrtc_sphere_cost.argtypes = [_vp, _vp, i32, i32, P(Robot), _vp, i32, f32, _vp, _vp, _vp]
    L.prrtc_sphere_cost.restype = ctypes.c_int
    L.prrtc_refine_workspace_bytes.argtypes = [i32, i32, i32]
    L.prrtc_refine_workspace_bytes.restype = sz
    L.prrtc_refine_paths.argtypes = [_vp, _vp, i32, i32, i32, _vp, _vp, P(Robot), P(World), i32, f32, f32, f32, f32,
                                     i32, _vp, _vp, sz, _vp]
    L.prrtc_refine_paths.restype = ctypes.c_int
    f3 = ctypes.POINTER(ctypes.c_float)
    L.prrtc_cloud_bytes.argtypes = [i32, f3, f3, f32]
    L.prrtc_cloud_bytes.restype = sz
    L.prrtc_cloud_build.argtypes = [_vp, _vp, i32, f3, f3, f32, f32, _vp, sz, P(Cloud)]
    L.prrtc_cloud_build.restype = ctypes.c_int
    L.prrtc_check_edges.argtypes = [_vp, _vp, i32, i32, i32, P(Robot), P(World), i32, f32, P(Cloud), _vp, P(Launch),
                                    _vp]
    L.prrtc_check_edges.restype = ctypes.c_int
    L.prrtc_last_error.argtypes = []
    L.prrtc_last_error.restype = ctypes.c_char_p
    return L


def library_source_hash(version: str) -> str:
    """The source hash a ``prrtc_version()`` string carries ("" if none)."""
    return version.rsplit(" src ", 1)[1].strip() if " src " in version else ""


def _check_fresh(version: str) -> None:
    from . import build

    if not build.source_files():
        return  # sources not shipped with this install: nothing to compare against
    want, have = build.source_hash(), library_source_hash(version)
    if have != want:
        raise RuntimeError(
            f"{LIB_PATH} was built from other sources (library {have or 'without a hash'}, sources {want}); "
            "rebuild it with `python -m prrax_amd.build`")


def check(fn: str, rc: int) -> None:
    if rc != PRRTC_OK:
        msg = lib().prrtc_last_error()
        raise PrrtcError(fn, rc, msg.decode() if msg else "")
