"""Pins the FP64 sphere-penetration cost and its adjoint gradient (the oracle of the refinement
kernels): closed form, central differences, torch float64 autograd, prismatic / screw / mimic joints."""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from prrax_amd import scenes
from prrax_amd.refine import resample_path

from helpers import torch_sphere_cost64, with_prismatic_and_screw


def _contexts():
    c5 = scenes.scene_c5().context()
    return {"c5": c5, "c4": scenes.scene_c4().context(), "c5_prism_screw": with_prismatic_and_screw(c5)}


def _near_obstacle_configs(cc, lo, hi, n, seed):
    """Configurations with a non-trivial cost (some spheres within a few cm of an obstacle)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        q = rng.uniform(lo, hi)
        if O.sphere_cost64(cc, q) > 1e-3:
            out.append(q)
    return np.array(out)


def test_cost_closed_form_single_pair():
    """One robot sphere at the tool of a 1-joint arm vs one world sphere: softplus by hand."""
    cc = {
        "fk_twists": np.array([[0, 0, 0, 0, 0, 1]], np.float32),
        "fk_parent_tf": np.array([[1, 0, 0, 0, 0, 0, 0]], np.float32),
        "fk_parent_idx": np.array([-1], np.int32), "fk_act_idx": np.array([0], np.int32),
        "fk_mimic_mul": np.ones(1, np.float32), "fk_mimic_off": np.zeros(1, np.float32),
        "fk_mimic_act_idx": np.array([-1], np.int32), "fk_topo_inv": np.array([0], np.int32),
        "sphere_link_idx": np.array([0], np.int32), "sphere_local": np.array([[0.5, 0, 0]], np.float32),
        "sphere_radius": np.array([0.1], np.float32), "world_spheres": np.array([[0.5, 0.2, 0, 0.15]], np.float32),
    }
    beta = 50.0
    L = float(np.float32(0.5))                       # the context holds float32 values
    o = np.array([L, float(np.float32(0.2)), 0.0])
    rr = float(np.float32(0.1)) + float(np.float32(0.15))
    for th in (0.0, 0.3, -1.0, 2.0):
        c = np.array([L * math.cos(th), L * math.sin(th), 0.0])
        d = np.linalg.norm(c - o)
        x = rr - d
        want = math.log1p(math.exp(beta * x)) / beta
        assert O.sphere_cost64(cc, np.array([th])) == pytest.approx(want, rel=1e-12)
        # d/dth: sigmoid(beta x) * (-(c - o)/d) . dc/dth, dc/dth = (-L sin, L cos, 0)
        sig = 1.0 / (1.0 + math.exp(-beta * x))
        dc = np.array([-L * math.sin(th), L * math.cos(th), 0.0])
        g = sig * (-(c - o) / d) @ dc
        assert O.sphere_cost_grad64(cc, np.array([th]))[0] == pytest.approx(g, rel=1e-10, abs=1e-14)


@pytest.mark.parametrize("name", ["c5", "c4", "c5_prism_screw"])
def test_adjoint_matches_central_differences(name):
    cc = _contexts()[name]
    robot = scenes.scene_c4().robot if name == "c4" else scenes.scene_c5().robot
    qs = _near_obstacle_configs(cc, robot.lower, robot.upper, 6, seed=7)
    h = 1e-4
    dim = qs.shape[1]
    for q in qs:
        g = O.sphere_cost_grad64(cc, q)
        fd = np.array([(O.sphere_cost64(cc, q + h * e) - O.sphere_cost64(cc, q - h * e)) / (2 * h)
                       for e in np.eye(dim)])
        assert np.abs(g - fd).max() <= 1e-5 * max(1.0, np.abs(fd).max()), (g, fd)


@pytest.mark.parametrize("name", ["c5", "c5_prism_screw"])
def test_adjoint_matches_torch_autograd(name):
    cc = _contexts()[name]
    robot = scenes.scene_c5().robot
    qs = _near_obstacle_configs(cc, robot.lower, robot.upper, 3, seed=11)
    for q in qs:
        qt = torch.tensor(q, dtype=torch.float64, requires_grad=True)
        c = torch_sphere_cost64(cc, qt)
        c.backward()
        assert float(c) == pytest.approx(O.sphere_cost64(cc, q), rel=1e-10)
        np.testing.assert_allclose(O.sphere_cost_grad64(cc, q), qt.grad.numpy(), rtol=1e-8, atol=1e-10)


def test_adjoint_catches_a_wrong_sign():
    """The central-difference check fails when the moment term's sign is flipped in a copy of
    the adjoint (a planted mistake in the oracle, on the CPU)."""
    cc = _contexts()["c5"]
    robot = scenes.scene_c5().robot
    q = _near_obstacle_configs(cc, robot.lower, robot.upper, 1, seed=3)[0]
    g = O.sphere_cost_grad64(cc, q)
    h = 1e-4
    fd = np.array([(O.sphere_cost64(cc, q + h * e) - O.sphere_cost64(cc, q - h * e)) / (2 * h) for e in np.eye(7)])
    assert np.abs(-g - fd).max() > 1e-3 * max(1.0, np.abs(fd).max())


def test_resample_path():
    p = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 2.0]], np.float32)
    r = resample_path(p, 7)
    assert r.shape == (7, 2)
    np.testing.assert_array_equal(r[0], p[0])
    np.testing.assert_array_equal(r[-1], p[-1])
    seg = np.linalg.norm(np.diff(r, axis=0), axis=1)
    np.testing.assert_allclose(seg, 0.5, atol=1e-6)  # total length 3 over 6 equal steps
    assert resample_path(p[:1], 4).shape == (4, 2)
