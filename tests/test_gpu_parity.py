"""GPU parity: every algorithm through the C ABI against the CPU oracle on the same
seeded inputs.  Integer outputs (BFS levels, SSSP distances, CC labels, BC levels
and path counts) must be bit-exact; BC dependencies and PageRank ranks are checked
against FP64 with the tolerance stated in each test."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMALL = {}


def graphs(mg):
    if SMALL:
        return SMALL
    SMALL["rmat16"] = mg.Graph.rmat(16, 16, seed=1)                     # BFS from vertex 0, RMAT-16
    SMALL["rmat12_sparse"] = mg.Graph.rmat(12, 2, seed=5)               # many components
    SMALL["grid64"] = mg.Graph.grid(64, 48)
    n = 300
    SMALL["path"] = mg.Graph.from_coo(n, np.arange(n - 1), np.arange(1, n), np.full(n - 1, 3))
    m = 20000                                                            # hub above the split threshold
    SMALL["star"] = mg.Graph.from_coo(m + 1, np.zeros(m, np.int32), np.arange(1, m + 1), np.arange(m) % 64 + 1)
    SMALL["edge"] = mg.Graph.from_coo(2, [0], [1], [5])
    return SMALL


def sources(g):
    nv = g.nv
    return sorted({0, g.max_degree_vertex(), nv // 3, nv - 1})


@pytest.fixture(scope="module", params=["rmat16", "rmat12_sparse", "grid64", "path", "star", "edge"])
def case(request, mg):
    g = graphs(mg)[request.param]
    row, col, w = g.csr()
    return request.param, g, row, col, w


MODES = [("hybrid", dict()), ("sparse", dict(mode=1)), ("dense", dict(mode=2)), ("graph", dict(graph_levels=3))]


@pytest.mark.parametrize("mode", MODES, ids=[m[0] for m in MODES])
def test_bfs(mg, orc, case, mode):
    name, g, row, col, _ = case
    e = mg.Engine(g, mg.BFS, **mode[1])
    for s in sources(g):
        print(f"bfs {name} mode={mode[0]} src={s}", flush=True)
        e.run(s)
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), orc.bfs(row, col, s))
    e.close()


def test_bfs_step_api(mg, orc, case):
    """the reference's `while (MultiGraph_V_E_V(...));` contract: nonzero = call again"""
    name, g, row, col, _ = case
    e = mg.Engine(g, mg.BFS)
    s = g.start_p
    e.reset(s)
    steps = 0
    while True:
        more, it = e.step()
        steps += 1
        if not more:
            break
    ref = orc.bfs(row, col, s)
    np.testing.assert_array_equal(e.result(mg.RES_LEVEL), ref)
    assert steps == int(ref[ref < orc.INF].max()) + 1
    e.close()


@pytest.mark.parametrize("graph_levels", [0, 4])
def test_sssp(mg, orc, case, graph_levels):
    name, g, row, col, w = case
    e = mg.Engine(g, mg.SSSP, graph_levels=graph_levels)
    for s in sources(g):
        print(f"sssp {name} src={s}", flush=True)
        e.run(s)
        ref, _ = orc.sssp_bf(row, col, w, s)
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), ref)
    e.close()


def test_cc(mg, orc, case):
    name, g, row, col, _ = case
    e = mg.Engine(g, mg.CC)
    e.run(0)
    lab = e.result(mg.RES_LEVEL)
    ref, k, _ = orc.cc_labelprop(row, col)
    uf, _ = orc.cc_unionfind(row, col)
    np.testing.assert_array_equal(lab, uf)            # min-label form, bit-exact
    np.testing.assert_array_equal(orc.canon(lab), ref)  # the reference's first-appearance form
    e.close()


@pytest.mark.parametrize("mode", MODES[:3], ids=[m[0] for m in MODES[:3]])
def test_bc(mg, orc, case, mode):
    name, g, row, col, _ = case
    e = mg.Engine(g, mg.BC, **mode[1])
    for s in sources(g)[:2]:
        print(f"bc {name} mode={mode[0]} src={s}", flush=True)
        e.run(s)
        lv, sg, dl = orc.bc_fp64(row, col, s)
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), lv)
        np.testing.assert_array_equal(e.result(mg.RES_SIGMA), sg)       # exact: integers below 2^53
        d = e.result(mg.RES_DELTA).astype(np.float64)
        # FP32 storage of delta: relative error within 1e-6 of max(1, |delta|)
        np.testing.assert_allclose(d, dl, rtol=1e-6, atol=1e-6)
    e.close()


@pytest.mark.parametrize("algo", ["PR_T", "PR_D"])
def test_pagerank_tight(mg, orc, case, algo):
    """tol 1e-7, at most 100 iterations: FP32 ranks vs the FP64 oracle, L1 <= 1e-6 relative"""
    name, g, row, col, _ = case
    a = mg.PR_T if algo == "PR_T" else mg.PR_D
    e = mg.Engine(g, a, pr_tolerance=1e-7, pr_max_iter=100)
    e.run(0)
    r = e.result(mg.RES_RANK).astype(np.float64)
    # The stopping rules compare each change with the tolerance, so FP32 and FP64 can
    # stop at different iterations on oscillating inputs (bipartite stars); the FP64
    # oracle is run for the iteration count the GPU took, which isolates arithmetic error.
    its = e.iterations()
    ref, _ = orc.pr_fp64(row, col, 0 if algo == "PR_T" else 1, 0.85, 1e-7, -its)
    l1 = np.abs(r - ref).sum() / max(np.abs(ref).sum(), 1e-300)
    print(f"pr {name} {algo} iters={its} L1rel={l1:.3e} maxabs={np.abs(r - ref).max():.3e}")
    assert l1 <= 1e-6
    e.close()


def test_pagerank_reference_rule(mg, orc, case):
    """the reference's own setting (tol 0.01, unbounded): same iteration count band and
    ranks within the tolerance the rule itself allows"""
    name, g, row, col, _ = case
    e = mg.Engine(g, mg.PR_T)
    e.run(0)
    r = e.result(mg.RES_RANK).astype(np.float64)
    ref, it = orc.pr_fp64(row, col, 0, 0.85, 0.01, 0)
    assert abs(e.iterations() - it) <= 2
    assert np.abs(r - ref).max() <= 0.2
    e.close()


@pytest.mark.parametrize("n", [300, 1000, 5000])
def test_cc_long_paths(mg, orc, n):
    """long chains stress the concurrent hooks and the final compression: every label
    must be its component's root, run after run"""
    g = mg.Graph.from_coo(n, np.arange(n - 1), np.arange(1, n), np.full(n - 1, 3))
    row, col, _ = g.csr()
    uf, _ = orc.cc_unionfind(row, col)
    e = mg.Engine(g, mg.CC)
    for _ in range(5):
        e.run(0)
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), uf)
    e.close()


@pytest.fixture(scope="module")
def directed(mg):
    """RMAT read without symmetrization: every level follows out-edges only"""
    g = mg.Graph.rmat(12, 8, seed=6, flags=0)
    row, col, w = g.csr()
    return g, row, col, w


@pytest.mark.parametrize("mode", MODES, ids=[m[0] for m in MODES])
def test_directed_bfs_sssp(mg, orc, directed, mode):
    g, row, col, w = directed
    eb = mg.Engine(g, mg.BFS, **mode[1])
    es = mg.Engine(g, mg.SSSP, **mode[1])
    for s in sources(g):
        print(f"directed bfs/sssp mode={mode[0]} src={s}", flush=True)
        eb.run(s)
        np.testing.assert_array_equal(eb.result(mg.RES_LEVEL), orc.bfs(row, col, s))
        es.run(s)
        np.testing.assert_array_equal(es.result(mg.RES_LEVEL), orc.sssp_bf(row, col, w, s)[0])
    eb.close()
    es.close()


@pytest.mark.parametrize("mode", MODES, ids=[m[0] for m in MODES])
def test_directed_bc(mg, orc, directed, mode):
    """pull levels of a directed graph read the transposed CSR (forward) and push the
    backward pass along in-edges"""
    g, row, col, w = directed
    e = mg.Engine(g, mg.BC, **mode[1])
    for s in sources(g):
        print(f"directed bc mode={mode[0]} src={s}", flush=True)
        e.run(s)
        lv, sg, dl = orc.bc_fp64(row, col, s)
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), lv)
        np.testing.assert_array_equal(e.result(mg.RES_SIGMA), sg)
        np.testing.assert_allclose(e.result(mg.RES_DELTA).astype(np.float64), dl, rtol=1e-6, atol=1e-6)
    e.close()


def fan_in(mg, n=5000):
    """0 -> 1..n -> n+1 -> n+2 (directed): the hub n+1 has n in-edges, above the pull
    split threshold, and one out-edge; weights vary per edge"""
    src = np.concatenate([np.zeros(n, np.int32), np.arange(1, n + 1, dtype=np.int32), [n + 1]]).astype(np.int32)
    dst = np.concatenate([np.arange(1, n + 1, dtype=np.int32), np.full(n, n + 1, np.int32), [n + 2]]).astype(np.int32)
    val = (1 + (np.arange(len(src)) * 7919) % 64).astype(np.int32)
    return mg.Graph.from_coo(n + 3, src, dst, val, flags=0)


@pytest.mark.parametrize("mode", MODES, ids=[m[0] for m in MODES])
def test_directed_fan_in(mg, orc, mode):
    g = fan_in(mg)
    row, col, w = g.csr()
    for algo in (mg.BFS, mg.SSSP, mg.BC):
        e = mg.Engine(g, algo, **mode[1])
        for s in (0, 1, g.nv - 2):
            print(f"fan-in algo={algo} mode={mode[0]} src={s}", flush=True)
            e.run(s)
            if algo == mg.BFS:
                np.testing.assert_array_equal(e.result(mg.RES_LEVEL), orc.bfs(row, col, s))
            elif algo == mg.SSSP:
                np.testing.assert_array_equal(e.result(mg.RES_LEVEL), orc.sssp_bf(row, col, w, s)[0])
            else:
                lv, sg, dl = orc.bc_fp64(row, col, s)
                np.testing.assert_array_equal(e.result(mg.RES_LEVEL), lv)
                np.testing.assert_array_equal(e.result(mg.RES_SIGMA), sg)
                np.testing.assert_allclose(e.result(mg.RES_DELTA).astype(np.float64), dl, rtol=1e-6, atol=1e-6)
        e.close()


def test_directed_cc_pr(mg, orc, directed):
    g, row, col, w = directed
    e = mg.Engine(g, mg.CC)
    e.run(0)
    np.testing.assert_array_equal(e.result(mg.RES_LEVEL), orc.cc_unionfind(row, col)[0])
    e.close()
    for mode in (dict(), dict(mode=1)):
        e = mg.Engine(g, mg.PR_T, pr_tolerance=1e-7, pr_max_iter=100, **mode)
        e.run(0)
        r = e.result(mg.RES_RANK).astype(np.float64)
        ref, _ = orc.pr_fp64(row, col, 0, 0.85, 1e-7, -e.iterations())
        assert np.abs(r - ref).sum() / np.abs(ref).sum() <= 1e-6
        e.close()


def test_from_csr_engines(mg, orc):
    """a caller-built CSR with shuffled rows runs like the reader's graph: BFS, SSSP (the
    layout sweeps need the rows sorted by column, which from_csr guarantees), BC and CC
    bit-exact against the oracle on the sorted CSR"""
    base = mg.Graph.rmat(12, 16, seed=8)
    row, col, w = base.csr()
    rng = np.random.default_rng(4)
    sc, sw = col.copy(), w.copy()
    for v in range(base.nv):
        p = rng.permutation(np.arange(row[v], row[v + 1]))
        sc[row[v]:row[v + 1]] = col[p]
        sw[row[v]:row[v + 1]] = w[p]
    g = mg.Graph.from_csr(row.astype(np.int64), sc, sw)
    r2, c2, w2 = g.csr()
    np.testing.assert_array_equal(c2, col)
    np.testing.assert_array_equal(w2, w)
    s = base.max_degree_vertex()
    for algo, ref in ((mg.BFS, orc.bfs(row, col, s)), (mg.SSSP, orc.sssp_bf(row, col, w, s)[0]),
                      (mg.BC, orc.bc_fp64(row, col, s)[0])):
        e = mg.Engine(g, algo, mode=2)
        e.run(s)
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), ref)
        e.close()
    e = mg.Engine(g, mg.CC)
    e.run(0)
    np.testing.assert_array_equal(e.result(mg.RES_LEVEL), orc.cc_unionfind(row, col)[0])
    e.close()


SHAPES = [
    dict(block_threads=512, lane_edges=4, pull_split=64, back_split=16, back_chunk=64),
    dict(block_threads=1024, lane_edges=200, pull_split=100000, back_split=100000, back_chunk=4096),
    dict(block_threads=256, lane_edges=1, pull_split=64, back_split=2, back_chunk=64, huge_degree=4096),
]


@pytest.mark.parametrize("shape", SHAPES, ids=["b512-fine", "b1024-coarse", "b256-finest"])
@pytest.mark.parametrize("gname", ["rmat16", "star", "grid64"])
def test_bin_edges_and_blocks(mg, orc, gname, shape):
    """Bin edges and workgroup sizes move work between the lane / wave / chunk tiers
    and change the grids, never the results: BFS and SSSP bit-exact, BC levels and
    path counts bit-exact, dependencies within the FP32 tolerance."""
    g = graphs(mg)[gname]
    row, col, w = g.csr()
    src = sources(g)[:2]
    for algo in (mg.BFS, mg.SSSP, mg.BC):
        for mode in (0, 2):  # hybrid, and every level a pull
            e = mg.Engine(g, algo, mode=mode, **shape)
            for s in src:
                print(f"{gname} algo={algo} mode={mode} {shape} src={s}", flush=True)
                e.run(s)
                got = e.result(mg.RES_LEVEL)
                if algo == mg.BFS:
                    np.testing.assert_array_equal(got, orc.bfs(row, col, s))
                elif algo == mg.SSSP:
                    np.testing.assert_array_equal(got, orc.sssp_bf(row, col, w, s)[0])
                else:
                    lv, sg, dl = orc.bc_fp64(row, col, s)
                    np.testing.assert_array_equal(got, lv)
                    np.testing.assert_array_equal(e.result(mg.RES_SIGMA), sg)
                    np.testing.assert_allclose(e.result(mg.RES_DELTA).astype(np.float64), dl, rtol=1e-6, atol=1e-6)
            e.close()


def test_block_threads_rejected(mg):
    g = graphs(mg)["edge"]
    with pytest.raises(mg.MGError):
        mg.Engine(g, mg.BFS, block_threads=300)


@pytest.mark.parametrize("mode", MODES, ids=[m[0] for m in MODES])
def test_wide_offsets(mg, orc, case, mode):
    """Engines with 64-bit row offsets (what a graph past 2^31 - 1 edges gets; forced
    here on small graphs) give the same BFS levels and SSSP distances as the oracle"""
    name, g, row, col, w = case
    if mode[0] == "sparse":
        with pytest.raises(mg.MGError):  # push-only levels need int32 frontier edge counts
            mg.Engine(g, mg.BFS, wide_offsets=1, **mode[1])
        return
    for algo in (mg.BFS, mg.SSSP):
        e = mg.Engine(g, algo, wide_offsets=1, **mode[1])
        for s in sources(g):
            print(f"wide {name} algo={algo} mode={mode[0]} src={s}", flush=True)
            e.run(s)
            ref = orc.bfs(row, col, s) if algo == mg.BFS else orc.sssp_bf(row, col, w, s)[0]
            np.testing.assert_array_equal(e.result(mg.RES_LEVEL), ref)
        e.close()


def test_wide_offsets_limits(mg):
    """64-bit row offsets serve BFS and SSSP on symmetric graphs; the rest is refused"""
    g = graphs(mg)["rmat16"]
    for algo in (mg.BC, mg.CC, mg.PR_T):
        with pytest.raises(mg.MGError):
            mg.Engine(g, algo, wide_offsets=1)
    with pytest.raises(mg.MGError):
        mg.Engine(g, mg.BFS, wide_offsets=1, alpha=2.0)
    d = mg.Graph.from_coo(4, [0, 1, 2], [1, 2, 3], flags=0)  # directed path
    assert not d.symmetric
    with pytest.raises(mg.MGError):
        mg.Engine(d, mg.BFS, wide_offsets=1)
