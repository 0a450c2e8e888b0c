"""prrtc_planner_noalloc (workspace in a static device array) against prrtc_planner (caller
workspace), through the C ABI, bit for bit: two calls in a row, the second with a smaller tree
capacity and fewer goals, so it runs on what the first call left in the shared array."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from prrax_amd import _lib, scenes
from prrax_amd.context import check_context, upload_robot, upload_world
from prrax_amd.prrtc import PlanBuffers, _attrs

from helpers import query

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 16 << 20  # PRRTC_NOALLOC_SCRATCH_BYTES


def _outputs(bufs: PlanBuffers) -> dict:
    h = bufs.host_summary()
    sa, sb = (int(v) for v in h["sizes"][0])
    n = max(int(h["path_len"][0]), 0)
    return {
        "ints": {k: np.asarray(h[k]).copy() for k in ("sizes", "completed", "iterations", "conn", "solved", "status",
                                                      "path_len")},
        "tree_a": bufs.tree_a[0, :, :sa].cpu().numpy().view(np.uint32).copy(),
        "tree_b": bufs.tree_b[0, :, :sb].cpu().numpy().view(np.uint32).copy(),
        "par_a": bufs.par_a[0, :sa].cpu().numpy().copy(),
        "par_b": bufs.par_b[0, :sb].cpu().numpy().copy(),
        "path": bufs.path[0, :n].cpu().numpy().view(np.uint32).copy(),
    }


def _run(entry: str, start, goals, robot, world, lo, hi, attrs, max_nodes: int, S: int) -> dict:
    dev = torch.device("cuda", 0)
    bufs = PlanBuffers(1, 7, max_nodes, S, 1024, dev)
    st = torch.from_numpy(np.ascontiguousarray(start, np.float32)).to(dev)
    gl = torch.from_numpy(np.ascontiguousarray(goals, np.float32)).to(dev)
    L = _lib.lib()
    stream = torch.cuda.current_stream().cuda_stream
    launch = _lib.Launch(0, 0, -1, 0, 0)
    if entry == "noalloc":
        rc = L.prrtc_planner_noalloc(stream, st.data_ptr(), gl.data_ptr(), int(goals.shape[0]), lo.data_ptr(),
                                     hi.data_ptr(), robot.struct, world.struct, attrs, bufs.outputs(), launch)
    else:
        rc = L.prrtc_planner(stream, st.data_ptr(), gl.data_ptr(), int(goals.shape[0]), lo.data_ptr(), hi.data_ptr(),
                             robot.struct, world.struct, attrs, bufs.outputs(), launch, bufs.workspace.data_ptr(),
                             bufs.workspace.numel())
    _lib.check(f"prrtc_planner ({entry})", rc)
    torch.cuda.synchronize()
    return _outputs(bufs)


def _same(a: dict, b: dict, tag: str) -> None:
    for k, v in a["ints"].items():
        np.testing.assert_array_equal(v, b["ints"][k], err_msg=f"{tag} {k}")
    for k in ("tree_a", "tree_b", "par_a", "par_b", "path"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{tag} {k}")


def test_noalloc_matches_planner_twice():
    torch.cuda.set_device(0)
    sc = scenes.scene_c1()
    cc = sc.context()
    arrays = check_context(cc, 7)
    dev = torch.device("cuda", 0)
    robot, world = upload_robot(arrays, dev), upload_world(arrays, dev)
    lo = torch.from_numpy(sc.robot.lower.astype(np.float32)).to(dev)
    hi = torch.from_numpy(sc.robot.upper.astype(np.float32)).to(dev)
    calls = [(3, 3, 16384, 128), (5, 1, 4096, 64)]  # (query, goals, max_nodes, S): the second is smaller
    for qi, ng, N, S in calls:
        start, goals = query(sc, 1, qi, n_goals=ng)
        attrs = _attrs(7, 10_000, 0.5, S, 16, N, 1, 0.5, True, 1e-4, 4.0, 1.0, 0, 0.02)
        print(f"noalloc vs planner: query {qi}, {ng} goals, max_nodes {N}, S {S}", flush=True)
        want = _run("planner", start, goals, robot, world, lo, hi, attrs, N, S)
        got = _run("noalloc", start, goals, robot, world, lo, hi, attrs, N, S)
        if N == calls[0][2]:
            assert int(want["ints"]["solved"][0]) == 1, "the first query should be solved"
        _same(got, want, f"query {qi}")


def test_noalloc_rejects_oversized_workspace():
    torch.cuda.set_device(0)
    sc = scenes.scene_c1()
    arrays = check_context(sc.context(), 7)
    dev = torch.device("cuda", 0)
    robot, world = upload_robot(arrays, dev), upload_world(arrays, dev)
    N, S = 4_000_000, 128
    L = _lib.lib()
    need = int(L.prrtc_planner_workspace_bytes(1, 7, N, S))
    assert need > SCRATCH_BYTES
    attrs = _attrs(7, 10, 0.5, S, 16, N, 1, 0.5, True, 1e-4, 4.0, 1.0, 0, 0.02)
    z = torch.zeros(64, dtype=torch.float32, device=dev)
    zi = torch.zeros(64, dtype=torch.int32, device=dev)
    out = _lib.Outputs(z.data_ptr(), z.data_ptr(), zi.data_ptr(), zi.data_ptr(), zi.data_ptr(), zi.data_ptr(),
                       zi.data_ptr(), zi.data_ptr(), zi.data_ptr(), z.data_ptr(), None, None, None, 0)
    rc = L.prrtc_planner_noalloc(torch.cuda.current_stream().cuda_stream, z.data_ptr(), z.data_ptr(), 1, z.data_ptr(),
                                 z.data_ptr(), robot.struct, world.struct, attrs, out, None)
    assert rc == _lib.PRRTC_E_WORKSPACE
    assert b"prrtc_planner_noalloc holds" in L.prrtc_last_error()
