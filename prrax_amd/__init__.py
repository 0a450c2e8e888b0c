"""prrax_amd -- MI355X-native pRRTC motion planner.

A drop-in for the ``prrtc_plan`` API of CoMMALab/prrax (cuda-rrtc/jax/__init__.py:5-14): the
planner, its FK and collision checks and its nearest-neighbour search run as hand-written HIP
kernels for gfx950 behind the C ABI in ``include/prrtc_amd.h``.
"""

from .prrtc import (PRRTCResult, CollisionChecker, forward_kinematics, prepare_batch, prrtc_extend,
                    prrtc_nearest_neighbor, prrtc_plan, prrtc_plan_batch)
from .robogpu import PointCloud, PointCloudChecker, prrtc_plan_robogpu

__version__ = "0.1.0"

__all__ = [
    "prrtc_plan",
    "prrtc_plan_batch",
    "prrtc_nearest_neighbor",
    "prrtc_extend",
    "PRRTCResult",
    "CollisionChecker",
    "forward_kinematics",
    "prepare_batch",
    "prrtc_plan_robogpu",
    "PointCloudChecker",
    "PointCloud",
]
