#!/usr/bin/env python3
"""Run one benchmark workload a few times, for a profiler wrapped around this process
(rocprofv3 --kernel-trace --stats / --pmc ... -- python3 multigraph_amd/workload_runs.py NAME).

    python3 multigraph_amd/workload_runs.py NAME [--reps K] [--warmup W]

NAME (graphs as bench.py builds them, multigraph_amd.workloads):
    sssp_cc22     SSSP from the highest-degree vertex + CC, RMAT-22 (the headline step)
    bfs22, bc22   BFS / BC from the highest-degree vertex, RMAT-22
    pr24_t, pr24_d  PageRank PR_T / PR_D, tolerance 1e-7, at most 100 iterations, RMAT-24
    batch_bfs22, batch_bc22, batch_sssp22   64-source batches (default_rng(3)), RMAT-22
    uneven_bfs, uneven_sssp                 the 32-graph uneven batch, default options
Each run prints its time (HIP events) on stdout.  No checking: the tests do that.
"""
import argparse
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import multigraph_amd as mg  # noqa: E402
from multigraph_amd import workloads as wl  # noqa: E402

NAMES = ("sssp_cc22", "bfs22", "bc22", "pr24_t", "pr24_d", "batch_bfs22", "batch_bc22", "batch_sssp22", "uneven_bfs",
         "uneven_sssp")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("name", choices=NAMES)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    a = ap.parse_args()
    t0 = time.time()
    n = a.name
    if n.startswith("uneven"):
        g, _, _, src, slots = wl.uneven(mg)
    else:
        g = wl.rmat(mg, 24 if n.startswith("pr24") else 22)
        src, slots = None, None
    print(f"[{n}] graph nv={g.nv} ne={g.ne} ({time.time() - t0:.1f}s)", flush=True)
    if n == "sssp_cc22":
        s = g.max_degree_vertex()
        e1, e2 = mg.Engine(g, mg.SSSP), mg.Engine(g, mg.CC)
        run = lambda: e1.run(s) + e2.run(0)
    elif n in ("bfs22", "bc22"):
        s = g.max_degree_vertex()
        e = mg.Engine(g, mg.BFS if n == "bfs22" else mg.BC)
        run = lambda: e.run(s)
    elif n.startswith("pr24"):
        e = mg.Engine(g, mg.PR_T if n == "pr24_t" else mg.PR_D, pr_tolerance=1e-7, pr_max_iter=100)
        run = lambda: e.run(0)
    else:
        kind = {"bfs": mg.BATCH_BFS, "bc": mg.BATCH_BC, "sssp": mg.BATCH_SSSP}[n.split("_")[1].replace("22", "")]
        b = mg.Batch(g, kind)
        if src is None:
            src = wl.batch_sources(g.nv)
        run = lambda: b.run(src, slots)
    for i in range(a.warmup + a.reps):
        ms = run()
        print(f"[{n}] {'warmup' if i < a.warmup else 'run'} {ms:.3f} ms", flush=True)


if __name__ == "__main__":
    main()
