"""Run with MG_LIBRARY pointing at multigraph_amd/libmgamd_hookvar.so (the engine built
with tests/hooks/variant_hooks.hpp).  Every path of every changed algorithm must show
the changed hook:
  SSSP (edges weigh 2w): distances are exactly twice the oracle's, for push levels,
    layout sweeps, CSR sweeps, pull-only runs and batched runs;
  BC (path-count messages doubled): levels as the oracle's, sigma[v] = 2^level(v) times
    the oracle's, for single-source runs in every mode and batched runs;
  PageRank (messages halved): ranks follow r = 0.15 + 0.85 * sum 0.5 r_u / deg_u,
    checked against that rule iterated in FP64 for the same number of iterations;
  BFS (hooks unchanged): levels as the oracle's.
Exits 0 when every check holds; prints the first failure otherwise."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import multigraph_amd as mg  # noqa: E402
from oracle import oracle as orc  # noqa: E402

INF = mg.INF


def say(msg):
    print(f"[hooks] {msg}", flush=True)


def pr_half(row, col, its, tol):
    """the halved-message rule, PR_T's update test, FP64, exactly `its` iterations"""
    nv = len(row) - 1
    deg = np.diff(row).astype(np.float64)
    src = np.repeat(np.arange(nv), np.diff(row))
    r = np.zeros(nv)
    for _ in range(its):
        c = np.where(deg > 0, 0.5 * r / np.maximum(deg, 1), 0.0)
        nr = np.bincount(col, weights=c[src], minlength=nv)
        f1 = 0.15 + 0.85 * nr
        upd = np.abs(f1 - r) > tol
        r = np.where(upd, f1, r)
    return r


def main():
    assert os.path.basename(mg.LIB_PATH) == "libmgamd_hookvar.so", mg.LIB_PATH
    graphs = {"rmat14": mg.Graph.rmat(14, 16, seed=2), "grid": mg.Graph.grid(40, 30)}
    for name, g in graphs.items():
        row, col, w = g.csr()
        s = g.max_degree_vertex()
        ref_d = orc.sssp_bf(row, col, w, s)[0].astype(np.int64)
        want = np.where(ref_d == INF, INF, 2 * ref_d)
        for opts in (dict(), dict(pull_layout=1), dict(pull_layout=2), dict(mode=mg.MODE_SPARSE), dict(mode=mg.MODE_DENSE)):
            e = mg.Engine(g, mg.SSSP, **opts)
            e.run(s)
            got = e.result(mg.RES_LEVEL)
            e.close()
            assert np.array_equal(got, want), f"SSSP {name} {opts}: not twice the oracle"
        say(f"SSSP {name}: distances doubled on every path")
        lv, sg, _ = orc.bc_fp64(row, col, s)
        want_sg = np.where(lv == INF, 0.0, sg * np.exp2(np.where(lv == INF, 0, lv)))
        for opts in (dict(), dict(mode=mg.MODE_SPARSE), dict(mode=mg.MODE_DENSE)):
            e = mg.Engine(g, mg.BC, **opts)
            e.run(s)
            assert np.array_equal(e.result(mg.RES_LEVEL), lv), f"BC levels {name} {opts}"
            assert np.array_equal(e.result(mg.RES_SIGMA), want_sg), f"BC sigma {name} {opts}"
            e.close()
        say(f"BC {name}: sigma scaled by 2^level on every path")
        e = mg.Engine(g, mg.BFS)
        e.run(s)
        assert np.array_equal(e.result(mg.RES_LEVEL), orc.bfs(row, col, s)), f"BFS {name}"
        e.close()
        src = np.random.default_rng(1).integers(0, g.nv, 16).astype(np.int32)
        b = mg.Batch(g, mg.BATCH_SSSP)
        b.run(src)
        got = b.levels().astype(np.int64)
        b.close()
        for i, x in enumerate(src):
            d = orc.sssp_dijkstra(row, col, w, int(x)).astype(np.int64)
            assert np.array_equal(got[i], np.where(d == INF, INF, 2 * d)), f"batched SSSP {name} row {i}"
        b = mg.Batch(g, mg.BATCH_BC)
        b.run(src)
        blv, bsg = b.levels(), b.sigma()
        b.close()
        for i, x in enumerate(src):
            l1, s1, _ = orc.bc_fp64(row, col, int(x))
            assert np.array_equal(blv[i], l1), f"batched BC levels {name} row {i}"
            assert np.array_equal(bsg[i], np.where(l1 == INF, 0.0, s1 * np.exp2(np.where(l1 == INF, 0, l1)))), \
                f"batched BC sigma {name} row {i}"
        say(f"batched SSSP / BC {name}: changed hooks seen in all 16 rows")
        for opts in (dict(), dict(mode=mg.MODE_SPARSE)):
            e = mg.Engine(g, mg.PR_T, pr_tolerance=1e-7, pr_max_iter=60, **opts)
            e.run(0)
            its = e.iterations()
            r = e.result(mg.RES_RANK).astype(np.float64)
            e.close()
            ref = pr_half(row, col, its, 1e-7)
            l1 = np.abs(r - ref).sum() / np.abs(ref).sum()
            normal, _ = orc.pr_fp64(row, col, 0, 0.85, 1e-7, -its)
            l1n = np.abs(r - normal).sum() / np.abs(normal).sum()
            assert l1 <= 1e-6, f"PageRank {name} {opts}: L1 {l1:.2e} against the halved rule"
            assert l1n > 1e-2, f"PageRank {name} {opts}: still the unchanged rule"
            say(f"PageRank {name} {opts}: halved messages (L1 {l1:.1e}; {l1n:.2f} from the unchanged rule)")
    say("every changed hook is called on every path")


if __name__ == "__main__":
    main()
