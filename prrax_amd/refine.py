"""Differentiable sphere-penetration cost through FK, and planner-path refinement.

No reference counterpart (the reference computes no gradients; the nearest code is its margin
predicate, prrtc_planner.cu:314-387).  The cost of a configuration q is

    C(q) = sum over robot spheres i, world spheres w of softplus_beta(r_i + r_w - |c_i(q) - o_w|)

with sphere centres c_i through FK.  ``SphereCost`` runs the forward and the hand-written adjoint
kernels of the library; ``sphere_cost`` wraps them as a torch autograd function whose backward
recomputes FK (the analogue of a ``jax.custom_vjp``); ``refine_paths`` runs whole gradient
descents on the device, keeping a moved waypoint only if the exact collision predicate passes.
"""

from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import _lib
from .context import as_numpy, check_context, upload_robot, upload_world
from .prrtc import _device, _limits, _resolve_collision_mode, _stream_ptr


class SphereCost:
    """Forward / adjoint kernels for one robot and the world spheres of a collision context."""

    def __init__(self, collision_context: dict, dim: int, beta: float = 50.0):
        if not beta > 0.0:
            raise ValueError("beta must be > 0")
        self.dim = int(dim)
        self.beta = float(beta)
        self.arrays = check_context(collision_context, self.dim)
        self.device = _device()
        self.robot = upload_robot(self.arrays, self.device)
        self.world = upload_world(self.arrays, self.device)
        self.ws = self.world.tensors["world_spheres"]
        self.n_ws = int(self.world.struct.n_world_spheres)

    def _configs(self, q) -> torch.Tensor:
        if isinstance(q, torch.Tensor) and q.is_cuda:
            return q.detach().to(torch.float32).contiguous().reshape(-1, self.dim)
        return torch.from_numpy(as_numpy(q, np.float32).reshape(-1, self.dim)).to(self.device)

    def __call__(self, q, grad: bool = False, flops: Optional[torch.Tensor] = None):
        """Cost ``[n]`` (and gradient ``[n, dim]`` when ``grad``) of configurations ``[n, dim]``."""
        qt = self._configs(q)
        n = int(qt.shape[0])
        cost = torch.empty(n, dtype=torch.float32, device=self.device)
        g = torch.empty((n, self.dim), dtype=torch.float32, device=self.device) if grad else None
        rc = _lib.lib().prrtc_sphere_cost(
            _stream_ptr(), qt.data_ptr(), n, self.dim, self.robot.struct,
            self.ws.data_ptr() if self.n_ws else None, self.n_ws, self.beta, cost.data_ptr(),
            None if g is None else g.data_ptr(), None if flops is None else flops.data_ptr())
        _lib.check("prrtc_sphere_cost", rc)
        return (cost, g) if grad else cost


class _SphereCostFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q: torch.Tensor, sc: SphereCost):  # noqa: D401 - autograd protocol
        ctx.sc = sc
        ctx.save_for_backward(q)
        return sc(q)

    @staticmethod
    def backward(ctx, gout: torch.Tensor):
        (q,) = ctx.saved_tensors
        _, g = ctx.sc(q, grad=True)  # FK recomputed: no per-link transforms were kept
        return (g * gout.reshape(-1, 1)).reshape(q.shape).to(q.dtype), None


def sphere_cost(q: torch.Tensor, sc: SphereCost) -> torch.Tensor:
    """Differentiable per-configuration cost ``[n]`` of ``q [n, dim]`` (a device tensor)."""
    return _SphereCostFn.apply(q, sc)


def resample_path(path, n: int = 64) -> np.ndarray:
    """``n`` waypoints evenly spaced by arc length along a piecewise-linear path (ends kept)."""
    p = as_numpy(path, np.float64)
    if p.ndim != 2 or p.shape[0] < 1:
        raise ValueError("path must be [L, dim] with L >= 1")
    if n < 2:
        raise ValueError("n must be >= 2")
    if p.shape[0] == 1:
        return np.repeat(p, n, axis=0).astype(np.float32)
    seg = np.linalg.norm(np.diff(p, axis=0), axis=1)
    s = np.concatenate([[0.0], np.cumsum(seg)])
    if s[-1] == 0.0:
        return np.repeat(p[:1], n, axis=0).astype(np.float32)
    t = np.linspace(0.0, s[-1], n)
    out = np.empty((n, p.shape[1]))
    for d in range(p.shape[1]):
        out[:, d] = np.interp(t, s, p[:, d])
    out[0], out[-1] = p[0], p[-1]
    return out.astype(np.float32)


def refine_paths(paths, collision_context: dict, iterations: int = 50, step: float = 1e-3, beta: float = 50.0,
                 smooth_weight: float = 0.1, collision_checker: str = "binary", collision_margin: float = 0.02,
                 min_vals=None, max_vals=None, flops: Optional[torch.Tensor] = None):
    """Gradient descent on ``paths [P, L, dim]`` (all L-waypoint paths; end points fixed).

    Cost = sum of ``sphere_cost`` over waypoints + ``smooth_weight`` * sum |q_{k+1} - q_k|^2.
    Each step moves every interior waypoint by ``-step * gradient`` (clamped to the limits) and
    keeps the move only where the exact predicate (``collision_checker``) finds it free.
    Returns ``(paths [P, L, dim] device tensor, costs [iterations + 1] numpy)``.
    """
    mode = _resolve_collision_mode(collision_checker)
    p0 = paths if isinstance(paths, torch.Tensor) else torch.from_numpy(as_numpy(paths, np.float32))
    if p0.ndim != 3:
        raise ValueError("paths must be [n_paths, n_waypoints, dim]")
    P, L, dim = (int(x) for x in p0.shape)
    if L < 2:
        raise ValueError("paths need at least 2 waypoints")
    if iterations < 0:
        raise ValueError("iterations must be >= 0")
    arrays = check_context(collision_context, dim)
    dev = _device()
    robot = upload_robot(arrays, dev)
    world = upload_world(arrays, dev)
    lo, hi = _limits(min_vals, max_vals, dim, dev)
    q = p0.to(device=dev, dtype=torch.float32).contiguous().clone()
    L_ = _lib.lib()
    ws = torch.empty(int(L_.prrtc_refine_workspace_bytes(P, L, dim)), dtype=torch.uint8, device=dev)
    costs = torch.empty(int(iterations) + 1, dtype=torch.float32, device=dev)
    rc = L_.prrtc_refine_paths(_stream_ptr(), q.data_ptr(), P, L, dim, lo.data_ptr(), hi.data_ptr(), robot.struct,
                               world.struct, mode, float(collision_margin), float(beta), float(smooth_weight),
                               float(step), int(iterations), costs.data_ptr(), ws.data_ptr(), ws.numel(),
                               None if flops is None else flops.data_ptr())
    _lib.check("prrtc_refine_paths", rc)
    return q, costs.cpu().numpy()
