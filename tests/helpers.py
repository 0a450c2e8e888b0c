"""Shared test helpers: scenes, oracle contexts and planner-result comparison."""

from __future__ import annotations

import numpy as np

from oracle import oracle as O
from prrax_amd import scenes


def oracle_free_fn(cc: dict, dim: int):
    ctx = O.Context(cc, dim, collision_mode=1, collision_margin=1e-5)
    return lambda q: O.min_distance(ctx, q) >= 1e-5


def query(scene: scenes.Scene, config: int, qi: int, n_goals: int = 1):
    cc = scene.context()
    rng = np.random.default_rng(scenes.query_seed(config, qi))
    sg = scenes.sample_free(rng, scene.robot, 1 + n_goals, oracle_free_fn(cc, scene.robot.dim))
    return sg[0], sg[1:]


def plan_digests(tree_a, tree_b, parents_a, parents_b, path) -> tuple:
    """SHA-256 digests (32 bytes each) of a plan's trees [dim, size] (FP32 bits), parents (int32)
    and path [L, dim] (FP32 bits; the digest of no bytes when there is none)."""
    import hashlib

    def h(a, dt):
        return hashlib.sha256(b"" if a is None else np.ascontiguousarray(np.asarray(a), dt).tobytes()).digest()

    return (h(tree_a, np.float32), h(tree_b, np.float32), h(parents_a, np.int32), h(parents_b, np.int32),
            h(path, np.float32))


def compare_plan(gpu, ora: dict, tag: str = "") -> None:
    """GPU PRRTCResult vs oracle plan dict, bit for bit."""
    assert gpu.status == ora["status"], f"{tag} status {gpu.status} != {ora['status']}"
    assert gpu.iterations == ora["iterations"], f"{tag} iterations {gpu.iterations} != {ora['iterations']}"
    assert (gpu.tree_a_size, gpu.tree_b_size) == tuple(ora["sizes"]), f"{tag} sizes"
    sa, sb = ora["sizes"]
    ta = gpu.tree_a_configs.cpu().numpy()
    tb = gpu.tree_b_configs.cpu().numpy()
    np.testing.assert_array_equal(ta.view(np.uint32), ora["tree_a"][:, :sa].view(np.uint32), err_msg=f"{tag} tree A")
    np.testing.assert_array_equal(tb.view(np.uint32), ora["tree_b"][:, :sb].view(np.uint32), err_msg=f"{tag} tree B")
    np.testing.assert_array_equal(gpu.tree_a_parents.cpu().numpy(), ora["parents_a"][:sa], err_msg=f"{tag} parents A")
    np.testing.assert_array_equal(gpu.tree_b_parents.cpu().numpy(), ora["parents_b"][:sb], err_msg=f"{tag} parents B")
    if ora["status"] == O.STATUS_SOLVED:
        np.testing.assert_array_equal(gpu.path.view(np.uint32), ora["path"].view(np.uint32), err_msg=f"{tag} path")


def torch_sphere_cost64(cc: dict, q, beta: float = 50.0):
    """The sphere-penetration cost written with torch float64 ops (FK by twist exponentials), so
    torch autograd gives an independent gradient.  Returns a 0-d tensor; q is a [dim] tensor."""
    import torch

    t64 = lambda a: torch.as_tensor(np.asarray(a, np.float64))  # noqa: E731
    tw, ptf = t64(cc["fk_twists"]), t64(cc["fk_parent_tf"])
    par, act, mim = (np.asarray(cc[k]) for k in ("fk_parent_idx", "fk_act_idx", "fk_mimic_act_idx"))
    mul, off = t64(cc["fk_mimic_mul"]), t64(cc["fk_mimic_off"])
    inv = np.asarray(cc["fk_topo_inv"])

    def skew(w):
        z = torch.zeros((), dtype=torch.float64)
        return torch.stack([torch.stack([z, -w[2], w[1]]), torch.stack([w[2], z, -w[0]]),
                            torch.stack([-w[1], w[0], z])])

    def quat(qq):
        w, x, y, z = qq / torch.linalg.norm(qq)
        return torch.stack([
            torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)]),
            torch.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)]),
            torch.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)])])

    n = par.shape[0]
    Rs, ts = [None] * n, [None] * n
    eye = torch.eye(3, dtype=torch.float64)
    for j in range(n):
        th = q[mim[j]] * mul[j] + off[j] if mim[j] >= 0 else (q[act[j]] if act[j] >= 0 else q.sum() * 0.0)
        v, w = tw[j, :3] * th, tw[j, 3:] * th
        a = torch.linalg.norm(tw[j, 3:]) * torch.abs(th)
        W = skew(w)
        if float(torch.linalg.norm(tw[j, 3:])) == 0.0:
            R, V = eye, eye
        else:
            A = torch.sin(a) / a
            B = (1 - torch.cos(a)) / a ** 2
            C = (a - torch.sin(a)) / a ** 3
            R = eye + A * W + B * (W @ W)
            V = eye + B * W + C * (W @ W)
        Rp, tp = quat(ptf[j, :4]), ptf[j, 4:]
        if par[j] >= 0:
            Rb, tb = Rs[par[j]] @ Rp, Rs[par[j]] @ tp + ts[par[j]]
        else:
            Rb, tb = Rp, tp
        Rs[j], ts[j] = Rb @ R, Rb @ (V @ v) + tb
    loc, rad = t64(cc["sphere_local"]), t64(cc["sphere_radius"])
    ws = t64(np.asarray(cc["world_spheres"], np.float64).reshape(-1, 4))
    total = q.sum() * 0.0
    for s, l in enumerate(np.asarray(cc["sphere_link_idx"])):
        if not (0 <= l < n):
            continue
        j = inv[l]
        c = Rs[j] @ loc[s] + ts[j]
        d = torch.linalg.norm(c[None, :] - ws[:, :3], dim=1)
        total = total + torch.nn.functional.softplus(beta * (rad[s] + ws[:, 3] - d)).sum() / beta
    return total


def with_prismatic_and_screw(cc: dict) -> dict:
    """A copy of an arm context whose joint 2 is prismatic along z and joint 4 a screw."""
    out = dict(cc)
    tw = np.array(cc["fk_twists"], np.float32, copy=True)
    tw[2] = [0.0, 0.0, 0.3, 0.0, 0.0, 0.0]
    tw[4] = [0.05, 0.0, 0.1, 0.0, 1.0, 0.0]
    out["fk_twists"] = tw
    return out
