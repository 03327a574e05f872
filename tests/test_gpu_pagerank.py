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
    }


@pytest.mark.parametrize("gname", ["rmat15", "rmat12x2", "grid", "star"])
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
