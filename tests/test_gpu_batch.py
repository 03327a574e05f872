"""Batched multi-source runs (one bit per source per vertex) against the oracle run
once per source: BFS levels and SSSP distances bit-exact, BC path counts bit-exact,
summed BC dependencies within FP32 tolerance of the FP64 Brandes sums."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hubs(mg):
    """Two hubs of 3000 and 1500 leaves joined by a path, plus leaf-to-leaf edges:
    hub rows span several 1024-edge chunks, whose partial results must merge."""
    rng = np.random.default_rng(21)
    src = [0] * 3000 + [1] * 1500 + [0]
    dst = list(range(2, 3002)) + list(range(3002, 4502)) + [1]
    a = rng.integers(2, 4502, size=4000)
    b = rng.integers(2, 4502, size=4000)
    keep = a != b
    src += a[keep].tolist()
    dst += b[keep].tolist()
    return mg.Graph.from_coo(4502, np.array(src), np.array(dst))


def _graphs(mg):
    return {
        "rmat13": mg.Graph.rmat(13, 16, seed=11),
        "rmat11x2": mg.Graph.rmat(11, 2, seed=12),
        "grid": mg.Graph.grid(30, 45),
        "hubs": _hubs(mg),
    }


GRAPHS = ["rmat13", "rmat11x2", "grid", "hubs"]


def _sources(g, k, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, g.nv, size=k).astype(np.int32)


MODES = [dict(), dict(mode=1), dict(mode=2)]


@pytest.mark.parametrize("gname", GRAPHS)
@pytest.mark.parametrize("k", [1, 7, 64])
@pytest.mark.parametrize("opts", MODES, ids=["hybrid", "push", "pull"])
def test_batch_bfs(mg, orc, gname, k, opts):
    g = _graphs(mg)[gname]
    row, col, _ = g.csr()
    src = _sources(g, k, 3 + k)
    b = mg.Batch(g, mg.BATCH_BFS, **opts)
    b.run(src)
    lv = b.levels()
    for i, s in enumerate(src):
        np.testing.assert_array_equal(lv[i], orc.bfs(row, col, int(s)), err_msg=f"source {i}={s}")
    b.close()


@pytest.mark.parametrize("gname", GRAPHS)
@pytest.mark.parametrize("opts", MODES, ids=["hybrid", "push", "pull"])
def test_batch_bc(mg, orc, gname, opts):
    g = _graphs(mg)[gname]
    row, col, _ = g.csr()
    src = _sources(g, 64, 5)
    src[7] = src[3]  # duplicate source: both slots must be served
    b = mg.Batch(g, mg.BATCH_BC, **opts)
    b.run(src)
    lv = b.levels()
    sg = b.sigma()
    bc = b.bc().astype(np.float64)
    ref_bc = np.zeros(g.nv)
    for i, s in enumerate(src):
        l, sig, dl = orc.bc_fp64(row, col, int(s))
        np.testing.assert_array_equal(lv[i], l)
        np.testing.assert_array_equal(sg[i], sig)
        dl = dl.copy()
        dl[int(s)] = 0.0
        ref_bc += dl
    err = np.abs(bc - ref_bc) / np.maximum(1.0, np.abs(ref_bc))
    print(f"bc {gname} max rel err {err.max():.2e}")
    assert err.max() <= 1e-5
    b.close()


@pytest.mark.parametrize("gname", GRAPHS)
def test_batch_sssp(mg, orc, gname):
    g = _graphs(mg)[gname]
    row, col, w = g.csr()
    src = _sources(g, 64, 9)
    b = mg.Batch(g, mg.BATCH_SSSP)
    b.run(src)
    d = b.levels()
    for i, s in enumerate(src):
        ref, _ = orc.sssp_bf(row, col, w, int(s))
        np.testing.assert_array_equal(d[i], ref, err_msg=f"source {i}={s}")
    b.close()


@pytest.mark.parametrize("opts", [dict(), dict(mode=2), dict(mode=2, sssp_delta=-1)], ids=["hybrid", "pull", "pull-bf"])
def test_batch_sssp_wide_distances(mg, orc, opts):
    """pull sweeps gather 16-bit copies of the distance rows until a finite distance
    above 65534 is met, then 32-bit rows for the rest of the run: weights up to 30000
    pass 65534 within a few hops, and a 1,500-vertex path of weight-64 edges does it
    after ~1,000 hops, mid-run; every row as the oracle's"""
    rng = np.random.default_rng(43)
    n, k = 2000, 1500
    src = np.concatenate([rng.integers(0, n, 16000), np.arange(n, n + k - 1), [0]])
    dst = np.concatenate([rng.integers(0, n, 16000), np.arange(n + 1, n + k), [n]])
    val = np.concatenate([rng.integers(1, 30001, 16000), np.full(k - 1, 64), [5]])
    g = mg.Graph.from_coo(n + k, src, dst, val)
    row, col, w = g.csr()
    sources = np.concatenate([[0, g.nv - 1, 7, n + 3], _sources(g, 60, 44)]).astype(np.int32)
    b = mg.Batch(g, mg.BATCH_SSSP, **opts)
    b.run(sources)
    d = b.levels()
    wide = False
    for i, s in enumerate(sources):
        ref, _ = orc.sssp_bf(row, col, w, int(s))
        wide |= bool(ref[ref != 0x64646464].max() > 65535)
        np.testing.assert_array_equal(d[i], ref, err_msg=f"{opts} source {i}={s}")
    assert wide
    b.close()


def _fan_in(mg, n=3000):
    """0 -> 1..n -> n+1 -> n+2, directed: the hub n+1 has n in-edges (several
    1024-edge in-chunks, merged by the pulls) and one out-edge"""
    src = np.concatenate([np.zeros(n, np.int32), np.arange(1, n + 1, dtype=np.int32), [n + 1]]).astype(np.int32)
    dst = np.concatenate([np.arange(1, n + 1, dtype=np.int32), np.full(n, n + 1, np.int32), [n + 2]]).astype(np.int32)
    val = (1 + (np.arange(len(src)) * 7919) % 64).astype(np.int32)
    return mg.Graph.from_coo(n + 3, src, dst, val, flags=0)


DIRECTED_MODES = [("hybrid", dict()), ("sparse", dict(mode=1)), ("dense", dict(mode=2))]


@pytest.mark.parametrize("mode", DIRECTED_MODES, ids=[m[0] for m in DIRECTED_MODES])
@pytest.mark.parametrize("gname", ["rmat", "fan_in"])
@pytest.mark.parametrize("kind", ["bfs", "sssp", "bc"])
def test_batch_directed(mg, orc, kind, gname, mode):
    """A directed CSR (no MG_READ_SYM): pushes follow out-edges, pulls read the
    transposed CSR; levels, distances and path counts as the oracle's per source"""
    g = mg.Graph.rmat(11, 8, seed=13, flags=0) if gname == "rmat" else _fan_in(mg)
    row, col, w = g.csr()
    src = _sources(g, 64, 17)
    if gname == "fan_in":
        src[:3] = [0, 1, g.nv - 2]
    b = mg.Batch(g, {"bfs": mg.BATCH_BFS, "sssp": mg.BATCH_SSSP, "bc": mg.BATCH_BC}[kind], **mode[1])
    print(f"directed batch {kind} {gname} mode={mode[0]}", flush=True)
    b.run(src)
    lv = b.levels()
    sg = b.sigma() if kind == "bc" else None
    for i, s in enumerate(src):
        if kind == "bfs":
            np.testing.assert_array_equal(lv[i], orc.bfs(row, col, int(s)))
        elif kind == "sssp":
            np.testing.assert_array_equal(lv[i], orc.sssp_bf(row, col, w, int(s))[0], err_msg=f"source {i}={s}")
        else:
            l, sig, _ = orc.bc_fp64(row, col, int(s))
            np.testing.assert_array_equal(lv[i], l, err_msg=f"source {i}={s}")
            np.testing.assert_array_equal(sg[i], sig, err_msg=f"source {i}={s}")
    b.close()


@pytest.mark.parametrize("mode", [dict(), dict(mode=1)], ids=["hybrid", "sparse"])
def test_batch_sssp_isolated_sources(mg, orc, mode):
    """vertices without edges (ids kept by from_csr, the last one included) as sources
    of queue levels: their changed bits relax nothing"""
    n = 40
    nbr = [[v - 1] * (v > 0) + [v + 1] * (v < n - 1) for v in range(n)] + [[], []]
    row = np.concatenate([[0], np.cumsum([len(x) for x in nbr])]).astype(np.int64)
    col = np.array([u for x in nbr for u in x], np.int32)
    u = np.repeat(np.arange(n + 2), np.diff(row))
    w = (1 + (np.minimum(u, col) * 13 + np.maximum(u, col)) % 7).astype(np.int32)  # same both ways
    g = mg.Graph.from_csr(row, col, w)
    row, col, w = g.csr()
    src = np.array([n, n + 1, 0, n // 2, n - 1, n + 1], np.int32)
    b = mg.Batch(g, mg.BATCH_SSSP, **mode)
    b.run(src)
    d = b.levels()
    for i, s in enumerate(src):
        np.testing.assert_array_equal(d[i], orc.sssp_bf(row, col, w, int(s))[0], err_msg=f"source {i}={s}")
    b.close()


@pytest.mark.parametrize("grid", [0, 7], ids=["grid-per-cu", "grid-7"])
@pytest.mark.parametrize("kind", ["bfs", "sssp", "bc"])
def test_batch_union_slots(mg, orc, kind, grid):
    """uneven batch: disjoint member graphs share the slots of one run (4 sources each);
    sparse levels on one workgroup per CU or on 7 (each workgroup then walks many queue
    entries)"""
    parts = [mg.Graph.rmat(10, 8, seed=31), mg.Graph.grid(23, 17), mg.Graph.rmat(8, 4, seed=32),
             mg.Graph.from_coo(50, np.arange(49), np.arange(1, 50), np.full(49, 2))]
    u = mg.Graph.union(parts)
    rng = np.random.default_rng(8)
    srcs, slots, offs = [], [], []
    off = 0
    for p in parts:
        offs.append(off)
        loc = rng.integers(0, p.nv, 4)
        srcs += (loc + off).tolist()
        slots += [0, 1, 2, 3]
        off += p.nv
    k = {"bfs": mg.BATCH_BFS, "sssp": mg.BATCH_SSSP, "bc": mg.BATCH_BC}[kind]
    b = mg.Batch(u, k, sparse_grid=grid)
    b.run(srcs, slots)
    lv = b.levels()
    assert lv.shape == (4, u.nv)
    bc = b.bc().astype(np.float64) if kind == "bc" else None
    for gi, p in enumerate(parts):
        row, col, w = p.csr()
        o = offs[gi]
        ref_bc = np.zeros(p.nv)
        for j in range(4):
            s = srcs[4 * gi + j] - o
            got = lv[j, o:o + p.nv]
            if kind == "sssp":
                np.testing.assert_array_equal(got, orc.sssp_bf(row, col, w, s)[0])
            else:
                np.testing.assert_array_equal(got, orc.bfs(row, col, s))
            if kind == "bc":
                _, _, dl = orc.bc_fp64(row, col, s)
                dl = dl.copy()
                dl[s] = 0.0
                ref_bc += dl
        if kind == "bc":
            err = np.abs(bc[o:o + p.nv] - ref_bc) / np.maximum(1.0, np.abs(ref_bc))
            assert err.max() <= 1e-5
    b.close()


@pytest.mark.parametrize("kind", ["bfs", "sssp"])
@pytest.mark.parametrize("graph_levels", [1, 3, 32])
def test_batch_graph_replay(mg, orc, kind, graph_levels):
    """levels replayed K at a time from a captured graph (past the end as no-ops), run
    twice on one batch object: every row matches the oracle, the level count does not
    depend on K"""
    g = mg.Graph.grid(40, 25)
    row, col, w = g.csr()
    k = mg.BATCH_BFS if kind == "bfs" else mg.BATCH_SSSP
    b = mg.Batch(g, k, graph_levels=graph_levels)
    ref_levels = None
    for seed in (1, 2):
        src = _sources(g, 64, 40 + seed)
        b.run(src)
        lv = b.levels()
        for i, s in enumerate(src):
            ref = orc.bfs(row, col, int(s)) if kind == "bfs" else orc.sssp_bf(row, col, w, int(s))[0]
            np.testing.assert_array_equal(lv[i], ref, err_msg=f"run {seed} source {i}={s}")
        if kind == "bfs":
            assert b.sweeps() == max(int(orc.bfs(row, col, int(s)).max()) for s in src) + 1
    b.close()


@pytest.mark.parametrize("graph_levels", [1, 8])
def test_batch_bfs_duplicate_appends(mg, orc, graph_levels):
    """Levels of the sparse-only graph queue a vertex once per discoverer that claims
    new bits.  Here 64 sources reach 64 middle vertices (one each), which all reach the
    same 16 vertices in one sparse level: 1,024 appends into a queue of nv + 1 = 465
    entries.  The appends past it are dropped, the next level rebuilds the queue from the
    frontier words, and a 20-vertex tail then runs on.  Every row matches the oracle."""
    src, dst = [], []
    for i in range(64):
        src.append(i); dst.append(64 + i)                      # source -> its middle vertex
        for j in range(16):
            src.append(64 + i); dst.append(128 + j)            # middle -> all 16 shared vertices
    f = np.arange(144, 444)
    a, c = np.triu_indices(len(f), 1)
    src += list(f[a]); dst += list(f[c])                       # a 300-vertex clique: E large, nv small
    src += [128] + list(range(444, 463)); dst += list(range(444, 464))  # a tail off the shared vertices
    g = mg.Graph.from_coo(464, np.array(src), np.array(dst))
    row, col, w = g.csr()
    assert g.nv == 464
    b = mg.Batch(g, mg.BATCH_BFS, graph_levels=graph_levels)
    sources = np.arange(64, dtype=np.int32)
    for rep in range(2):
        b.run(sources)
        lv = b.levels()
        for i in range(64):
            np.testing.assert_array_equal(lv[i], orc.bfs(row, col, i), err_msg=f"run {rep} source {i}")
    b.close()


@pytest.mark.parametrize("kind", ["bfs", "sssp"])
@pytest.mark.parametrize("persist", [0, 1, 7, -1])
def test_batch_persistent_levels(mg, orc, kind, persist):
    """sparse levels run by one cooperative launch (at most `persist` levels per launch,
    -1 = unbounded) or from the captured graphs only (0): every slot matches the oracle
    and the level count does not depend on the path.  Members: a long grid tail, a thin
    grid, and an RMAT graph whose hub forces chunked dense levels in between."""
    parts = [mg.Graph.grid(120, 90), mg.Graph.rmat(11, 8, seed=5), mg.Graph.grid(7, 300)]
    u = mg.Graph.union(parts)
    row, col, w = u.csr()
    rng = np.random.default_rng(19)
    srcs, slots, off = [], [], 0
    for p in parts:
        srcs += (rng.integers(0, p.nv, 3) + off).tolist()
        slots += [0, 1, 2]
        off += p.nv
    k = mg.BATCH_BFS if kind == "bfs" else mg.BATCH_SSSP
    b = mg.Batch(u, k, persist_levels=persist)
    sweeps = []
    persistent = []
    for rep in range(2):  # twice on one batch object
        b.run(srcs, slots)
        lv = b.levels()
        for j in range(3):
            ref = None
            for s, sl in zip(srcs, slots):
                if sl != j:
                    continue
                r = orc.bfs(row, col, s) if kind == "bfs" else orc.sssp_bf(row, col, w, s)[0]
                ref = r if ref is None else np.minimum(ref, r)
            np.testing.assert_array_equal(lv[j], ref, err_msg=f"run {rep} slot {j}")
        sweeps.append(b.sweeps())
        st = b.stats()
        assert st["levels"] == sweeps[-1] == st["dense_levels"] + st["sparse_levels"]
        persistent.append(st["persistent_levels"])
    b.close()
    assert sweeps[0] == sweeps[1]
    if persist == 0:
        assert persistent == [0, 0]
    else:
        assert min(persistent) > 0  # the grid tails ran in persistent launches
    if kind == "bfs":
        assert sweeps[0] == max(int(orc.bfs(row, col, s)[orc.bfs(row, col, s) != 0x64646464].max()) for s in srcs) + 1


def test_batch_persistent_large_sparse_levels(mg, orc):
    """persistent BFS levels run on one XCD only while a level's vertices + edges stay
    small: an RMAT-16 member whose wide levels are made sparse (low alpha and beta)
    sends those back to the sparse-only graph, and a 7 x 300 grid member's long tail
    then runs in persistent launches; every slot as the oracle's, the same level count
    as graph replay"""
    parts = [mg.Graph.rmat(16, 16, seed=27), mg.Graph.grid(7, 300)]
    u = mg.Graph.union(parts)
    row, col, _ = u.csr()
    rng = np.random.default_rng(28)
    src = np.concatenate([rng.integers(0, parts[0].nv, 32), parts[0].nv + rng.integers(0, parts[1].nv, 32)]).astype(np.int32)
    ref = orc.bfs_multi(row, col, src)
    runs = {}
    for persist in (0, -1):
        b = mg.Batch(u, mg.BATCH_BFS, persist_levels=persist, alpha=2, beta=1.5)
        b.run(src)
        np.testing.assert_array_equal(b.levels(), ref, err_msg=f"persist {persist}")
        runs[persist] = b.stats()
        b.close()
    print(runs, flush=True)
    assert runs[0]["levels"] == runs[-1]["levels"]
    assert runs[0]["persistent_levels"] == 0
    # the grid's tail ran persistent, the RMAT member's wide sparse levels did not
    assert 0 < runs[-1]["persistent_levels"] < runs[-1]["sparse_levels"]


@pytest.mark.parametrize("delta", [-1, 0, 1, 14, 16, 100000])
@pytest.mark.parametrize("gname", ["grid", "rmat13", "union"])
@pytest.mark.parametrize("persist", [0, -1], ids=["graph", "persistent"])
def test_batch_sssp_near_far(mg, orc, delta, gname, persist):
    """sparse sweeps relax only pairs below a rising distance threshold (the rest wait in
    the next changed word): distances bit-exact for any band, graph-replayed or
    persistent levels"""
    if gname == "union":
        parts = [mg.Graph.grid(60, 45), mg.Graph.rmat(10, 8, seed=31), mg.Graph.grid(5, 200)]
        g = mg.Graph.union(parts)
    else:
        g = _graphs(mg)[gname]
    row, col, w = g.csr()
    src = _sources(g, 64, 23)
    b = mg.Batch(g, mg.BATCH_SSSP, sssp_delta=delta, persist_levels=persist)
    b.run(src)
    d = b.levels()
    for i, s in enumerate(src):
        np.testing.assert_array_equal(d[i], orc.sssp_bf(row, col, w, int(s))[0], err_msg=f"delta {delta} source {i}={s}")
    b.close()


@pytest.mark.parametrize("kind", ["bfs", "sssp", "bc"])
def test_batch_lane_order(mg, orc, kind):
    """Sources placed in lanes by 2-hop reach (the default) give the rows of lanes in
    the caller's order: levels, distances and path counts equal, BC sums within FP32
    rounding, and every row is the oracle's for its source."""
    g = mg.Graph.rmat(13, 16, seed=11)
    row, col, w = g.csr()
    src = _sources(g, 64, 17)
    src[0] = g.max_degree_vertex()  # the widest reach lands in lane 0, so the order moves it
    k = {"bfs": mg.BATCH_BFS, "sssp": mg.BATCH_SSSP, "bc": mg.BATCH_BC}[kind]
    out = []
    for lanes in (0, 1):
        b = mg.Batch(g, k, batch_lanes=lanes)
        b.run(src)
        out.append((b.levels(), b.sigma() if kind == "bc" else None, b.bc() if kind == "bc" else None))
        b.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for i in (0, 5, 63):
        s = int(src[i])
        ref = orc.bfs(row, col, s) if kind != "sssp" else orc.sssp_dijkstra(row, col, w, s)
        np.testing.assert_array_equal(out[0][0][i], ref, err_msg=f"source {i}={s}")
    if kind == "bc":
        np.testing.assert_array_equal(out[0][1], out[1][1])
        a, b2 = out[0][2].astype(np.float64), out[1][2].astype(np.float64)
        assert np.max(np.abs(a - b2) / np.maximum(1.0, np.abs(b2))) < 1e-5
