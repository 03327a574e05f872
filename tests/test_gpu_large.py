"""Parity at a realistic size: symmetrized RMAT-20 (a=.57 b=.19 c=.19, edge factor 16,
seed 1; about 1 M vertices and 31 M directed edges after compaction), where the
partitioned layout has 256 partitions of ~3.9 K ids and hub rows span many chunks.
BFS levels, SSSP distances (layout and CSR sweeps), CC labels, BC levels and path
counts are bit-exact against the oracle; BC dependencies within 1e-6 relative of
max(1, |delta|); PageRank within 1e-6 relative L1 of FP64 run for the same number of
iterations.  Sources: the highest-degree vertex and a fixed random one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_G = {}


@pytest.fixture(scope="module")
def big(mg):
    if "g" not in _G:
        g = mg.Graph.rmat(20, 16, 0.57, 0.19, 0.19, seed=1, weight_seed=7)
        row, col, w = g.csr()
        srcs = [g.max_degree_vertex(), int(np.random.default_rng(11).integers(0, g.nv))]
        _G["g"] = (g, row, col, w, srcs)
    return _G["g"]


@pytest.mark.parametrize("opts", [dict(), dict(pull_layout=1), dict(mode=2)], ids=["layout", "csr", "pull-only"])
def test_large_sssp(mg, orc, big, opts):
    g, row, col, w, srcs = big
    e = mg.Engine(g, mg.SSSP, **opts)
    for s in srcs:
        print(f"sssp rmat20 {opts} src={s}", flush=True)
        e.run(s)
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), orc.sssp_bf(row, col, w, s)[0])
    e.close()


def test_large_bfs_cc(mg, orc, big):
    g, row, col, w, srcs = big
    e = mg.Engine(g, mg.BFS)
    for s in srcs:
        print(f"bfs rmat20 src={s}", flush=True)
        e.run(s)
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), orc.bfs(row, col, s))
    e.close()
    e = mg.Engine(g, mg.CC)
    e.run(0)
    np.testing.assert_array_equal(e.result(mg.RES_LEVEL), orc.cc_unionfind(row, col)[0])
    e.close()


def test_large_bc(mg, orc, big):
    g, row, col, w, srcs = big
    e = mg.Engine(g, mg.BC)
    for s in srcs:
        print(f"bc rmat20 src={s}", flush=True)
        e.run(s)
        lv, sg, dl = orc.bc_fp64(row, col, s)
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), lv)
        np.testing.assert_array_equal(e.result(mg.RES_SIGMA), sg)
        np.testing.assert_allclose(e.result(mg.RES_DELTA).astype(np.float64), dl, rtol=1e-6, atol=1e-6)
    e.close()


@pytest.mark.parametrize("algo", ["PR_T", "PR_D"])
def test_large_pagerank(mg, orc, big, algo):
    g, row, col, w, srcs = big
    a = mg.PR_T if algo == "PR_T" else mg.PR_D
    e = mg.Engine(g, a, pr_tolerance=1e-7, pr_max_iter=100)
    e.run(0)
    its = e.iterations()
    r = e.result(mg.RES_RANK).astype(np.float64)
    ref, _ = orc.pr_fp64(row, col, 0 if algo == "PR_T" else 1, 0.85, 1e-7, -its)
    l1 = np.abs(r - ref).sum() / np.abs(ref).sum()
    print(f"pr rmat20 {algo} iters={its} L1rel={l1:.3e} maxabs={np.abs(r - ref).max():.3e}")
    assert l1 <= 1e-6
    e.close()


def test_large_batched(mg, orc, big):
    """64-source batches (BFS, BC) against the oracle per source for 8 of the rows"""
    g, row, col, w, srcs = big
    src = np.random.default_rng(3).integers(0, g.nv, size=64).astype(np.int32)
    for kind in (mg.BATCH_BFS, mg.BATCH_BC):
        b = mg.Batch(g, kind)
        b.run(src)
        lv = b.levels()
        sg = b.sigma() if kind == mg.BATCH_BC else None
        for i in range(0, 64, 8):
            s = int(src[i])
            if kind == mg.BATCH_BFS:
                np.testing.assert_array_equal(lv[i], orc.bfs(row, col, s), err_msg=f"bfs row {i}")
            else:
                l, sig, _ = orc.bc_fp64(row, col, s)
                np.testing.assert_array_equal(lv[i], l, err_msg=f"bc row {i}")
                np.testing.assert_array_equal(sg[i], sig, err_msg=f"bc sigma row {i}")
        b.close()


def _bfs_sssp_at_scale(mg, orc, g, tag, offsets):
    """BFS levels bit-exact against the oracle; SSSP distances proven exact by their
    certificate (source 0, no edge lowers a distance, every reached vertex has a tight
    in-edge, reached set = BFS's); for each row-offset width in `offsets`"""
    import time
    t0 = time.time()
    row, col, w = g.csr(copy=False)
    src = g.max_degree_vertex()
    print(f"{tag} nv={g.nv} ne={g.ne} row {row.dtype} src={src}", flush=True)
    ref = orc.bfs(row, col, src)
    print(f"{tag} oracle bfs ({time.time() - t0:.0f} s)", flush=True)
    for wide in offsets:
        e = mg.Engine(g, mg.BFS, wide_offsets=wide)
        ms = e.run(src)
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), ref)
        e.close()
        print(f"{tag} wide_offsets={wide} bfs exact ({ms:.2f} ms)", flush=True)
        e = mg.Engine(g, mg.SSSP, wide_offsets=wide)
        ms = e.run(src)
        d = e.result(mg.RES_LEVEL)
        e.close()
        assert orc.sssp_certificate(row, col, w, src, d, ref) == 0
        print(f"{tag} wide_offsets={wide} sssp certificate holds ({ms:.2f} ms, {time.time() - t0:.0f} s)", flush=True)


def test_rmat26_bfs_sssp(mg, orc):
    """Symmetrized RMAT-26 (32.8 M vertices, 2.10 G directed edges after compaction,
    just inside int32 offsets; 2.15 G entries before deduplication), run with int32 row
    offsets and with 64-bit ones forced (the kernels of a graph past 2^31 - 1 edges)."""
    g = mg.Graph.rmat(26, 16, 0.57, 0.19, 0.19, seed=1, weight_seed=7)
    assert g.ne <= 2**31 - 1
    _bfs_sssp_at_scale(mg, orc, g, "rmat26", (0, 1))


@pytest.mark.skipif(not __import__("os").environ.get("MG_TEST_RMAT27"),
                    reason="several minutes and ~120 GB of host memory: set MG_TEST_RMAT27=1")
def test_rmat27_wide_bfs_sssp(mg, orc):
    """Past int32 offsets: symmetrized RMAT-27 (about 4.2 G directed edges), whose row
    offsets are 64-bit on the host (mg_graph_csr64) and on the device."""
    g = mg.Graph.rmat(27, 16, 0.57, 0.19, 0.19, seed=1, weight_seed=7)
    assert g.ne > 2**31 - 1
    _bfs_sssp_at_scale(mg, orc, g, "rmat27", (0,))
