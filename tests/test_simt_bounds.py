"""CPU build of the planner's wave-level collision checks, run under AddressSanitizer with bounds
assertions on their LDS queues and packed entries (tests/simt).

The device functions stage_model, fk_lane, wave_min_distance, wave_in_collision and
wave_unit_in_collision are compiled with g++ against a host stand-in for the HIP header that runs
one wave as 64 threads.  Every refinement-queue push (at most 2 * 64 entries), candidate-list push
(at most 2 * 64), packed entry (obstacle, configuration, group) and decoded index is asserted in
range, every cross-lane operation is checked to be reached by all 64 lanes at the same line, and
the work-group's LDS is one heap block of the model layout's exact size (ASan reports any access
past it).  The worlds cover the large scene, the 1024-query workload's worlds, degenerate
geometry (negative radii, skewed boxes, non-unit normals, zero-length capsules) on a tree robot,
the 14-DOF dual arm and a 64-frame chain (64 culling groups, the packing limit), with edges of
nearby configurations checked as sub-units.  Verdicts are compared with the oracle as on the GPU.
"""

from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from prrax_amd import scenes
from prrax_amd.context import check_context

from test_oracle_fk_collision import random_tree_robot

ROOT = Path(__file__).resolve().parent.parent
SIMT = ROOT / "tests" / "simt"
KFB, KUNIT_CAP, KCULL_Q = 16, 128, 128


@pytest.fixture(scope="module")
def simt_exe(tmp_path_factory) -> Path:
    exe = tmp_path_factory.mktemp("simt") / "simt_check"
    cmd = ["g++", "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-ffp-contract=off", "-Wall", "-Wno-attributes", "-Wno-unknown-pragmas",
           "-I", str(SIMT), "-I", str(ROOT / "prrax_amd" / "csrc"), "-I", str(ROOT / "include"),
           str(SIMT / "simt_check.cpp"), "-o", str(exe), "-lpthread"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def _write(path: Path, arrs: list) -> None:
    with open(path, "wb") as f:
        for a in arrs:
            a = np.ascontiguousarray(a).reshape(-1)
            f.write(np.int64(a.size).tobytes())
            f.write(a.tobytes())


def run_simt(exe: Path, tmp: Path, cc: dict, dim: int, q: np.ndarray, units: np.ndarray, mode: int,
             margin: float) -> dict:
    A = check_context(cc, dim)
    q = np.ascontiguousarray(q, np.float32)
    n = q.shape[0]
    units = np.ascontiguousarray(units, np.int32).reshape(-1, 2)
    inp, out = tmp / "in.bin", tmp / "out.bin"
    _write(inp, [A["fk_twists"], A["fk_parent_tf"], A["fk_parent_idx"], A["fk_act_idx"], A["fk_mimic_mul"],
                 A["fk_mimic_off"], A["fk_mimic_act_idx"], A["fk_topo_inv"], A["sphere_link_idx"], A["sphere_local"],
                 A["sphere_radius"], A["world_spheres"], A["world_capsules"], A["world_boxes"],
                 A["world_halfspaces"], A["self_pairs"], np.array([dim, mode, n, units.shape[0]], np.int32),
                 np.array([margin], np.float32), q, units])
    r = subprocess.run([str(exe), str(inp), str(out)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, f"simt_check failed ({r.returncode}):\n{r.stderr[-4000:]}"
    raw = out.read_bytes()
    nu = units.shape[0]
    o = 0
    hit_culled = np.frombuffer(raw, np.uint8, n, o); o += n
    md = np.frombuffer(raw, np.float32, n, o); o += 4 * n
    uv = np.frombuffer(raw, np.int32, nu, o); o += 4 * nu
    obs = np.frombuffer(raw, np.int64, 6, o)
    return {"hit_culled": hit_culled.astype(bool), "min": md, "unit": uv,
            "max_queue": int(obs[0]), "max_candidates": int(obs[1]), "max_unit_queue": int(obs[2]),
            "groups": int(obs[3]), "max_self_survivors": int(obs[4]), "self_overflows": int(obs[5])}


def edge_configs(rng, lo, hi, n_edges: int, step: float) -> np.ndarray:
    """n_edges batches of 16 waypoints t = (k + 1) / 16 along short random edges (the planner's
    discretisation), so consecutive configurations are close as in the planner's sub-units."""
    a = rng.uniform(lo, hi, size=(n_edges, lo.shape[0]))
    d = rng.normal(size=a.shape)
    d *= (step / np.linalg.norm(d, axis=1, keepdims=True))
    t = (np.arange(KFB) + 1.0) / KFB
    return (a[:, None, :] + t[None, :, None] * d[:, None, :]).reshape(-1, lo.shape[0]).astype(np.float32)


def unit_list(n: int) -> np.ndarray:
    """Sub-units of sizes 16, 8 + 8, 5 + 5 + 6, 3 + 13, 1 + 15, ... over consecutive batches."""
    splits = ([16], [8, 8], [5, 5, 6], [3, 13], [1, 15], [2, 2, 2, 10], [11, 5])
    out = []
    for b in range(n // KFB):
        c = b * KFB
        for s in splits[b % len(splits)]:
            out.append((c, s))
            c += s
    return np.array(out, np.int32)


def check_case(exe, tmp, cc, dim, q, mode, margin, expect_hits=(0.02, 0.98)):
    units = unit_list(q.shape[0])
    res = run_simt(exe, tmp, cc, dim, q, units, mode, margin)
    ctx = O.Context(cc, dim, collision_mode=mode, collision_margin=margin)
    want = O.in_collision(ctx, q)
    np.testing.assert_array_equal(res["hit_culled"], want)
    md = O.min_distance(O.Context(cc, dim, collision_mode=1, collision_margin=0.0), q)
    np.testing.assert_allclose(res["min"], md, rtol=0, atol=1e-6)
    lo_h, hi_h = expect_hits
    assert lo_h <= want.mean() <= hi_h, want.mean()
    # sub-unit verdicts: 0 iff every configuration is free; k + 1 names a colliding one; -1 only
    # when the candidate list overflowed
    n_over = 0
    for (c0, nu), v in zip(units, res["unit"]):
        part = want[c0:c0 + nu]
        if v == -1:
            n_over += 1
        elif v == 0:
            assert not part.any(), (c0, nu)
        else:
            assert 1 <= v <= nu and part[v - 1], (c0, nu, v)
    assert n_over < len(units)
    assert res["max_queue"] <= KCULL_Q and res["max_candidates"] <= KUNIT_CAP and res["max_unit_queue"] <= 64
    return res, n_over


@pytest.mark.parametrize("mode", [0, 1])
def test_simt_large_scene(simt_exe, tmp_path, mode):
    sc = scenes.scene_c2()
    rng = np.random.default_rng(5)
    q = np.concatenate([edge_configs(rng, sc.robot.lower, sc.robot.upper, 16, 0.05),
                        rng.uniform(sc.robot.lower, sc.robot.upper, size=(128, 7)).astype(np.float32)])
    cc = sc.context()
    res, _ = check_case(simt_exe, tmp_path, cc, 7, q, mode, 0.005)
    links = np.asarray(cc["sphere_link_idx"])
    assert res["groups"] == len(set(links[links >= 0].tolist()))  # one group per frame with spheres


def test_simt_workload_worlds(simt_exe, tmp_path):
    robot = scenes.make_arm(0, 64, 200)
    rng = np.random.default_rng(6)
    for qi in (0, 455, 777):
        w, _, _ = scenes.c3_query_world(robot, qi)
        cc = dict(robot.ctx, **w)
        q = edge_configs(rng, robot.lower, robot.upper, 12, 0.1)
        check_case(simt_exe, tmp_path, cc, 7, q, 0, 0.0, expect_hits=(0.0, 1.0))


@pytest.mark.parametrize("mode", [0, 1])
def test_simt_degenerate_geometry(simt_exe, tmp_path, mode):
    cc, dim = random_tree_robot(3)
    rng = np.random.default_rng(21)
    ws = np.concatenate([rng.uniform(-0.6, 0.6, (60, 3)), rng.uniform(-0.05, 0.12, (60, 1))], 1)
    caps = np.concatenate([rng.uniform(-0.6, 0.6, (6, 6)), rng.uniform(-0.02, 0.05, (6, 1))], 1)
    caps[0, 3:6] = caps[0, :3]
    boxes = np.zeros((6, 15))
    boxes[:, :3] = rng.uniform(-0.6, 0.6, (6, 3))
    for k in range(6):
        A = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        if k % 2:
            A = A * rng.uniform(0.5, 1.5, size=3) + 0.1 * rng.normal(size=(3, 3))
        boxes[k, 3:12] = A.T.reshape(-1)
        boxes[k, 12:] = rng.uniform(-0.02, 0.2, 3)
    hs = np.array([[0, 0, 1.3, 0, 0, -0.9], [0.2, 0.9, 0.1, 0, -1.2, 0]], np.float64)
    cc.update(world_spheres=ws.astype(np.float32), world_capsules=caps.astype(np.float32),
              world_boxes=boxes.astype(np.float32), world_halfspaces=hs.astype(np.float32))
    lo, hi = np.full(dim, -2.8), np.full(dim, 2.8)
    q = np.concatenate([edge_configs(rng, lo, hi, 16, 0.05), rng.uniform(lo, hi, size=(64, dim)).astype(np.float32)])
    check_case(simt_exe, tmp_path, cc, dim, q, mode, 0.01)


def test_simt_dual_arm(simt_exe, tmp_path):
    sc = scenes.scene_c4()
    rng = np.random.default_rng(8)
    q = edge_configs(rng, sc.robot.lower, sc.robot.upper, 8, 0.05)
    check_case(simt_exe, tmp_path, sc.context(), sc.robot.dim, q, 0, 0.0, expect_hits=(0.0, 1.0))


def chain64(rng, n_world: int) -> tuple[dict, int]:
    """64 frames in a chain (16 revolute joints, 48 fixed), two spheres per frame, and a dense
    cloud of small world spheres around the arm: 64 culling groups and long candidate lists."""
    nj, dim = 64, 16
    tw = np.zeros((nj, 6), np.float32)
    act = np.full(nj, -1, np.int32)
    for j in range(0, nj, 4):
        tw[j, 3 + (j // 4) % 3] = 1.0
        act[j] = j // 4
    ptf = np.zeros((nj, 7), np.float32)
    ptf[:, 0] = 1.0
    ptf[1:, 6] = 0.03
    links = np.repeat(np.arange(nj), 2).astype(np.int32)
    cc = {"fk_twists": tw, "fk_parent_tf": ptf, "fk_parent_idx": np.arange(-1, nj - 1).astype(np.int32),
          "fk_act_idx": act, "fk_mimic_mul": np.ones(nj, np.float32), "fk_mimic_off": np.zeros(nj, np.float32),
          "fk_mimic_act_idx": np.full(nj, -1, np.int32), "fk_topo_inv": np.arange(nj).astype(np.int32),
          "sphere_link_idx": links,
          "sphere_local": rng.uniform(-0.01, 0.01, size=(2 * nj, 3)).astype(np.float32),
          "sphere_radius": np.full(2 * nj, 0.012, np.float32),
          "self_pairs": np.array([[i, i + 40] for i in range(0, 2 * nj - 40, 3)], np.int32)}
    cen = rng.uniform(-1.0, 1.0, size=(n_world, 3))
    cc["world_spheres"] = np.concatenate([cen, rng.uniform(0.005, 0.02, (n_world, 1))], 1).astype(np.float32)
    cc["world_halfspaces"] = np.array([[0, 0, -2.5, 0, 0, -1]], np.float32)
    return cc, dim


def test_simt_64_groups_and_overflow(simt_exe, tmp_path):
    rng = np.random.default_rng(9)
    cc, dim = chain64(rng, 2000)
    lo, hi = np.full(dim, -0.6), np.full(dim, 0.6)
    q = np.concatenate([edge_configs(rng, lo, hi, 6, 0.02), edge_configs(rng, lo, hi, 4, 0.8)])
    res, n_over = check_case(simt_exe, tmp_path, cc, dim, q, 0, 0.0, expect_hits=(0.0, 1.0))
    assert res["groups"] == 64
    # long edges in a dense cloud fill the candidate list: the overflow path is exercised
    assert n_over > 0 and res["max_candidates"] > 64


@pytest.mark.parametrize("mode,message", [("ok", None), ("diverge", "divergent collective"),
                                          ("skip", "only 32 of 64 lanes arrived"),
                                          ("overflow", "assertion failed"), ("heap", "heap-buffer-overflow")])
def test_simt_harness_catches_planted_mistakes(tmp_path, mode, message):
    """The stand-in itself: lockstep collectives agree with wave64 semantics, and a divergent or
    skipped collective, a queue overflow and an access past an LDS-sized heap block all fail."""
    exe = tmp_path / "selftest"
    subprocess.run(["g++", "-std=c++20", "-O1", "-g", "-fsanitize=address", "-Wno-attributes", "-I", str(SIMT),
                    "-I", str(ROOT / "prrax_amd" / "csrc"), str(SIMT / "harness_selftest.cpp"), "-o", str(exe),
                    "-lpthread"], check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe), mode], capture_output=True, text=True, timeout=60)
    if message is None:
        assert r.returncode == 0, r.stderr
    else:
        assert r.returncode != 0 and message in r.stderr, r.stderr[-2000:]


@pytest.fixture(scope="module")
def simt_cloud_exe(tmp_path_factory) -> Path:
    exe = tmp_path_factory.mktemp("simtc") / "simt_cloud"
    cmd = ["g++", "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-ffp-contract=off", "-Wall", "-Wno-attributes", "-Wno-unknown-pragmas",
           "-I", str(SIMT), "-I", str(ROOT / "prrax_amd" / "csrc"), "-I", str(ROOT / "include"),
           str(SIMT / "simt_cloud.cpp"), "-o", str(exe), "-lpthread"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def run_simt_cloud(exe, tmp, cc, dim, q, points, r_env, lo, cell, shape, mode=0, margin=0.0):
    A = check_context(cc, dim)
    q = np.ascontiguousarray(q, np.float32)
    n = q.shape[0]
    inp, out = tmp / "cin.bin", tmp / "cout.bin"
    _write(inp, [A["fk_twists"], A["fk_parent_tf"], A["fk_parent_idx"], A["fk_act_idx"], A["fk_mimic_mul"],
                 A["fk_mimic_off"], A["fk_mimic_act_idx"], A["fk_topo_inv"], A["sphere_link_idx"], A["sphere_local"],
                 A["sphere_radius"], A["world_spheres"], A["world_capsules"], A["world_boxes"],
                 A["world_halfspaces"], A["self_pairs"], np.array([dim, mode, n, *shape], np.int32),
                 np.array([margin, *lo, cell, r_env], np.float32), np.asarray(points, np.float32), q])
    r = subprocess.run([str(exe), str(inp), str(out)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, f"simt_cloud failed ({r.returncode}):\n{r.stderr[-4000:]}"
    raw = out.read_bytes()
    return np.frombuffer(raw, np.uint8, n, 0).astype(bool), np.frombuffer(raw, np.uint8, n, n).astype(bool)


@pytest.mark.parametrize("grid", ["fit", "half", "fine", "one_cell", "clipped"])
def test_simt_cloud_grid_and_verdicts(simt_cloud_exe, tmp_path, grid):
    """wave_cloud_hit against a host-built grid: the same verdict as the oracle's brute force over
    every point, for cells of twice and half the largest reach (the default), cells much smaller
    than a reach (many rows per query), a single cell, and bounds smaller than the cloud (points
    clamped into border cells)."""
    sc = scenes.scene_cloud(6000)
    pts, r_env = sc.extra["points"], sc.extra["r_env"]
    cc = sc.context()
    rng = np.random.default_rng(12)
    q = np.concatenate([edge_configs(rng, sc.robot.lower, sc.robot.upper, 10, 0.05),
                        rng.uniform(sc.robot.lower, sc.robot.upper, size=(64, 7)).astype(np.float32)])
    lo, hi = pts.min(0), pts.max(0)
    reach = float(np.max(cc["sphere_radius"])) + r_env
    cell = {"fit": 2 * reach, "half": 0.5 * reach, "fine": 0.3 * reach, "one_cell": float((hi - lo).max()) + 1.0,
            "clipped": 2 * reach}[grid]
    if grid == "clipped":
        lo, hi = lo + 0.3 * (hi - lo), hi - 0.3 * (hi - lo)
    shape = [max(1, int(np.ceil((h - l) / cell))) for l, h in zip(lo, hi)]
    hw, hc = run_simt_cloud(simt_cloud_exe, tmp_path, cc, 7, q, pts, r_env, lo, cell, shape)
    np.testing.assert_array_equal(hw, O.in_collision(O.Context(cc, 7), q))
    want = O.cloud_hits(O.Context(cc, 7), q, pts, r_env)
    np.testing.assert_array_equal(hc, want)
    assert 0.05 < want.mean() < 0.95


@pytest.mark.parametrize("robot", ["arm", "dual"])
def test_simt_self_pairs_only(simt_exe, tmp_path, robot):
    """No world: every unit verdict comes from the self-pair stage (group-pair culling, the
    survivor list, and for the dual arm's 2000 pairs the fall-back over all pairs)."""
    sc = scenes.scene_c2() if robot == "arm" else scenes.scene_c4()
    cc = dict(sc.robot.ctx)
    dim = sc.robot.dim
    rng = np.random.default_rng(31)
    q = np.concatenate([edge_configs(rng, sc.robot.lower, sc.robot.upper, 12, 0.1),
                        rng.uniform(sc.robot.lower, sc.robot.upper, size=(64, dim)).astype(np.float32)])
    hits = (0.02, 0.98) if robot == "arm" else (0.0, 1.0)  # the two arms overlap in most configurations
    res, _ = check_case(simt_exe, tmp_path, cc, dim, q, 0, 0.0, expect_hits=hits)
    print(robot, "survivors", res["max_self_survivors"], "overflows", res["self_overflows"])
    check_case(simt_exe, tmp_path, cc, dim, q, 1, 0.01, expect_hits=hits)
