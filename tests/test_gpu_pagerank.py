"""PageRank on both device paths (partitioned layout and CSR pull), PR_T and PR_D,
against the FP64 oracle run for the same number of iterations (relative L1 <= 1e-6),
and the partition-size option across its range."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graphs(mg):
    return {
        "rmat15": mg.Graph.rmat(15, 16, seed=2),
        "rmat12x2": mg.Graph.rmat(12, 2, seed=4),
        "grid": mg.Graph.grid(33, 77),
        "star": mg.Graph.from_coo(9001, np.zeros(9000, np.int32), np.arange(1, 9001)),
        # the hub's partition receives 70000 edges: its gather is split over several
        # units and applied (and its messages written) by the separate passes
        "bigstar": mg.Graph.from_coo(70001, np.concatenate([np.zeros(70000, np.int32), np.arange(1, 70000)]),
                                     np.concatenate([np.arange(1, 70001), np.arange(2, 70001)])),
    }


@pytest.mark.parametrize("gname", ["rmat15", "rmat12x2", "grid", "star", "bigstar"])
@pytest.mark.parametrize("path", ["layout", "pull"])
@pytest.mark.parametrize("algo", ["PR_T", "PR_D"])
def test_pr_paths(mg, orc, gname, path, algo):
    g = _graphs(mg)[gname]
    row, col, _ = g.csr()
    a = mg.PR_T if algo == "PR_T" else mg.PR_D
    opts = dict(pr_tolerance=1e-7, pr_max_iter=60)
    if path == "pull":
        opts["mode"] = mg.MODE_SPARSE
    e = mg.Engine(g, a, **opts)
    e.run(0)
    its = e.iterations()
    r = e.result(mg.RES_RANK).astype(np.float64)
    ref, _ = orc.pr_fp64(row, col, 0 if algo == "PR_T" else 1, 0.85, 1e-7, -its)
    l1 = np.abs(r - ref).sum() / np.abs(ref).sum()
    print(f"{gname} {path} {algo} iters={its} L1rel={l1:.2e} maxabs={np.abs(r - ref).max():.2e}")
    assert l1 <= 1e-6
    e.close()


@pytest.mark.parametrize("plog", [8, 10, 12, 13, 14])
def test_pr_partition_sizes(mg, orc, plog):
    g = mg.Graph.rmat(14, 8, seed=6)
    row, col, _ = g.csr()
    e = mg.Engine(g, mg.PR_D, pr_tolerance=1e-7, pr_max_iter=30, part_log2=plog)
    e.run(0)
    its = e.iterations()
    ref, _ = orc.pr_fp64(row, col, 1, 0.85, 1e-7, -its)
    r = e.result(mg.RES_RANK).astype(np.float64)
    assert np.abs(r - ref).sum() / np.abs(ref).sum() <= 1e-6
    e.close()


@pytest.mark.parametrize("graph_levels", [1, 3, 8])
def test_pr_graph_replay_matches_steps(mg, orc, graph_levels):
    """iterations replayed K at a time from a captured graph (run past convergence
    as no-ops) give the same ranks and iteration count as stepping one at a time"""
    g = mg.Graph.rmat(13, 8, seed=9)
    e = mg.Engine(g, mg.PR_T, pr_tolerance=1e-5, pr_max_iter=0, graph_levels=graph_levels)
    e.run(0)
    r1, it1 = e.result(mg.RES_RANK), e.iterations()
    e.reset(0)
    steps = 0
    while True:
        more, _ = e.step()
        steps += 1
        if not more:
            break
    r2, it2 = e.result(mg.RES_RANK), e.iterations()
    # FP64 LDS sums are added in a scheduling-dependent order, so the two runs agree to
    # float rounding, not bit for bit
    np.testing.assert_allclose(r1, r2, rtol=1e-6)
    assert it2 == steps and abs(it1 - it2) <= 1
    row, col, _ = g.csr()
    ref, _ = orc.pr_fp64(row, col, 0, 0.85, 1e-5, -it1)
    assert np.abs(r1 - ref).sum() / np.abs(ref).sum() <= 1e-6
    e.close()
