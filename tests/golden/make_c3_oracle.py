"""Writes ``c3_s512_oracle.npz``: the oracle's plans of a sample of the uneven 1024-query
workload (7-DOF arm with 64 spheres and 200 self pairs; per query 10-300 world spheres, 0-24
boxes, 1-8 goals, max_nodes 8192-65536) at S = 512 samples per round, G = 16, step 0.5,
10,000 iterations: every 32nd query, q26 (52 rounds) and the two longest, q239 and q455 (they
fill their trees without connecting).

These are outputs of this repository's own oracle (``oracle/prrtc_oracle.c``, a restatement of
the reference kernel), stored because the long queries take it minutes: the GPU test
``tests/test_gpu_c3_full.py`` runs the whole 1024-query batch at the same settings and compares
each sampled query with these records.  Per query the file holds its inputs (start, goals,
max_nodes) and the plan: status, iterations, tree sizes and SHA-256 digests of both trees'
configurations and parents and of the path (``tests/helpers.plan_digests``).

    python tests/golden/make_c3_oracle.py      # about 6 minutes on 8 cores
"""
from __future__ import annotations

import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

QIDS = list(range(0, 1024, 32)) + [26, 239, 455]
MAX_GOALS = 8


def settings(robot) -> dict:
    return dict(max_iterations=10_000, step_size=0.5, num_new_samples=512, granularity=16, min_vals=robot.lower,
                max_vals=robot.upper)


def one(qi: int):
    from helpers import plan_digests, query
    from oracle import oracle as O
    from prrax_amd import scenes

    robot = scenes.make_arm(0, 64, 200)
    w, ng, N = scenes.c3_query_world(robot, qi)
    cc = dict(robot.ctx)
    cc.update(w)
    s, g = query(scenes.Scene(robot, w), 3, qi, n_goals=ng)
    r = O.plan(O.Context(cc, 7), s, g, max_nodes=N, **settings(robot))
    sa, sb = r["sizes"]
    d = plan_digests(r["tree_a"][:, :sa], r["tree_b"][:, :sb], r["parents_a"][:sa], r["parents_b"][:sb], r["path"])
    print(f"q{qi}: status {r['status']} iterations {r['iterations']} sizes {sa} {sb}", flush=True)
    return qi, s, g, N, r["status"], r["iterations"], (sa, sb), d


def main() -> None:
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        # longest first, so that they overlap the short ones
        rows = pool.map(one, sorted(QIDS, key=lambda q: q not in (239, 455)), chunksize=1)
    rows.sort(key=lambda r: r[0])
    K = len(rows)
    goals = np.zeros((K, MAX_GOALS, 7), np.float32)
    n_goals = np.zeros(K, np.int32)
    for k, r in enumerate(rows):
        n_goals[k] = r[2].shape[0]
        goals[k, :n_goals[k]] = r[2]
    np.savez_compressed(os.path.join(HERE, "c3_s512_oracle.npz"),
                        qids=np.array([r[0] for r in rows], np.int32),
                        starts=np.stack([r[1] for r in rows]).astype(np.float32), goals=goals, n_goals=n_goals,
                        max_nodes=np.array([r[3] for r in rows], np.int32),
                        status=np.array([r[4] for r in rows], np.int32),
                        iterations=np.array([r[5] for r in rows], np.int32),
                        sizes=np.array([r[6] for r in rows], np.int32),
                        digests=np.stack([np.frombuffer(b"".join(r[7]), np.uint8) for r in rows]))


if __name__ == "__main__":
    main()
