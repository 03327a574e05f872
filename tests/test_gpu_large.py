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


def test_rmat25_bfs_sssp(mg, orc):
    """The largest RMAT whose symmetrized CSR fits int32 offsets: scale 25 (17.1 M
    vertices, 1.05 G directed edges after compaction; hub rows over 800 K edges).
    BFS levels bit-exact against the oracle; SSSP distances proven exact by their
    certificate (d[source] = 0; d[v] <= d[u] + w on every edge; every other reached
    vertex has a tight in-edge; reached set = BFS's), checked in edge blocks."""
    g = mg.Graph.rmat(25, 16, 0.57, 0.19, 0.19, seed=1, weight_seed=7)
    row, col, w = g.csr()
    src = g.max_degree_vertex()
    print(f"rmat25 nv={g.nv} ne={g.ne} src={src}", flush=True)
    e = mg.Engine(g, mg.BFS)
    e.run(src)
    np.testing.assert_array_equal(e.result(mg.RES_LEVEL), orc.bfs(row, col, src))
    e.close()
    print("rmat25 bfs exact", flush=True)
    e = mg.Engine(g, mg.SSSP)
    e.run(src)
    d = e.result(mg.RES_LEVEL).astype(np.int64)
    e.close()
    inf = 0x64646464
    reached = d != inf
    assert d[src] == 0
    np.testing.assert_array_equal(reached, np.asarray(orc.bfs(row, col, src)) != inf)
    tight = np.zeros(g.nv, bool)
    blk = 1 << 26
    for p0 in range(0, g.ne, blk):
        p1 = min(g.ne, p0 + blk)
        u = np.searchsorted(row, np.arange(p0, p1), side="right") - 1
        du, dv, wp = d[u], d[col[p0:p1]], w[p0:p1]
        ok = du != inf
        assert np.all(dv[ok] <= du[ok] + wp[ok])
        tight[col[p0:p1][ok & (dv == du + wp)]] = True
    tight[src] = True
    assert np.all(tight[reached])
    print("rmat25 sssp certificate holds", flush=True)
