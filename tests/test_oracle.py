"""Pins the CPU oracle (oracle/oracle.c) independently of itself: networkx 3.4 and
scipy 1.15 routines, closed forms (grid Manhattan distances and binomial path
counts, paths, stars, regular graphs) and hand-iterated examples.  Each check is
chosen so that a plausible slip in the oracle (a dropped term, a wrong index or
direction, an off-by-one level) fails it; the planted-mistake tests at the end show
that the checks reject such slips.  CPU only."""
from math import comb

import networkx as nx
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse import csgraph

INF = 0x64646464


def csr_of(G, n=None):
    n = G.number_of_nodes() if n is None else n
    A = nx.to_scipy_sparse_array(G, nodelist=range(n), format="csr", weight=None)
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32)


def weights_of(G, row, col, key="w"):
    w = np.empty(len(col), np.int32)
    for u in range(len(row) - 1):
        for p in range(row[u], row[u + 1]):
            w[p] = G[u][int(col[p])][key]
    return w


def graphs():
    out = {
        "ba": nx.barabasi_albert_graph(300, 3, seed=1),
        "gnp": nx.gnp_random_graph(250, 0.02, seed=2),           # several components
        "grid": nx.convert_node_labels_to_integers(nx.grid_2d_graph(13, 9), ordering="sorted"),
        "path": nx.path_graph(40),
        "star": nx.star_graph(50),
        "complete": nx.complete_graph(12),
        "directed": nx.gnp_random_graph(200, 0.03, seed=3, directed=True),
    }
    rng = np.random.default_rng(7)
    for G in out.values():
        for u, v in G.edges():
            G[u][v]["w"] = int(rng.integers(1, 65))
    return out


GRAPHS = graphs()
NAMES = list(GRAPHS)


def _sources(n):
    return sorted({0, n // 2, n - 1})


# ---- BFS ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", NAMES)
def test_bfs_vs_networkx(orc, name):
    G = GRAPHS[name]
    row, col = csr_of(G)
    for s in _sources(G.number_of_nodes()):
        ref = np.full(G.number_of_nodes(), INF, np.int64)
        for v, d in nx.single_source_shortest_path_length(G, s).items():
            ref[v] = d
        np.testing.assert_array_equal(orc.bfs(row, col, s), ref)


def test_bfs_grid_manhattan(orc):
    W, H = 13, 9
    row, col = csr_of(GRAPHS["grid"])
    lv = orc.bfs(row, col, 0)                    # node (x, y) -> x * H + y (sorted labels)
    for x in range(W):
        for y in range(H):
            assert lv[x * H + y] == x + y


# ---- SSSP --------------------------------------------------------------------------
@pytest.mark.parametrize("name", NAMES)
def test_sssp_vs_scipy_and_dijkstra(orc, name):
    G = GRAPHS[name]
    row, col = csr_of(G)
    w = weights_of(G, row, col)
    n = len(row) - 1
    M = sp.csr_matrix((w.astype(np.float64), col, row), shape=(n, n))
    for s in _sources(n):
        bf, sweeps = orc.sssp_bf(row, col, w, s)
        dj = orc.sssp_dijkstra(row, col, w, s)
        ref = csgraph.dijkstra(M, directed=True, indices=s)
        ref = np.where(np.isinf(ref), INF, ref).astype(np.int64)
        np.testing.assert_array_equal(bf, ref)
        np.testing.assert_array_equal(dj, ref)
        assert sweeps >= 1
        # unit weights reduce to BFS
        np.testing.assert_array_equal(orc.sssp_bf(row, col, np.ones_like(w), s)[0], orc.bfs(row, col, s))


# ---- CC ----------------------------------------------------------------------------
@pytest.mark.parametrize("name", [n for n in NAMES if n != "directed"])
def test_cc_vs_scipy(orc, name):
    G = GRAPHS[name]
    row, col = csr_of(G)
    n = len(row) - 1
    k, lab = csgraph.connected_components(sp.csr_matrix((np.ones(len(col)), col, row), shape=(n, n)), directed=False)
    minid = np.full(k, n, np.int64)
    for v in range(n):
        minid[lab[v]] = min(minid[lab[v]], v)
    uf, kc = orc.cc_unionfind(row, col)
    np.testing.assert_array_equal(uf, minid[lab])
    assert kc == k
    lp, kl, _ = orc.cc_labelprop(row, col)
    assert kl == k
    np.testing.assert_array_equal(lp, orc.canon(minid[lab]))


def test_cc_directed_is_weak(orc):
    G = GRAPHS["directed"]
    row, col = csr_of(G)
    n = len(row) - 1
    k, lab = csgraph.connected_components(sp.csr_matrix((np.ones(len(col)), col, row), shape=(n, n)), connection="weak")
    uf, kc = orc.cc_unionfind(row, col)
    assert kc == k
    np.testing.assert_array_equal(orc.canon(uf), orc.canon(lab.astype(np.int32)))


def test_canon_hand_example(orc):
    np.testing.assert_array_equal(orc.canon([5, 5, 2, 2, 0, 5, 6]), [0, 0, 1, 1, 2, 0, 3])


# ---- BC ----------------------------------------------------------------------------
@pytest.mark.parametrize("name", NAMES)
def test_bc_vs_networkx(orc, name):
    G = GRAPHS[name]
    row, col = csr_of(G)
    n = len(row) - 1
    half = 1.0 if G.is_directed() else 0.5       # networkx halves undirected pair counts
    for s in _sources(n):
        lv, sg, dl = orc.bc_fp64(row, col, s)
        np.testing.assert_array_equal(lv, orc.bfs(row, col, s))
        # sigma: shortest paths enumerated by networkx, for targets with few of them
        reach = [v for v in range(n) if lv[v] != INF and v != s and sg[v] <= 500]
        for t in reach[:: max(1, len(reach) // 12)]:
            assert sg[t] == sum(1 for _ in nx.all_shortest_paths(G, s, t))
        ref = nx.betweenness_centrality_subset(G, sources=[s], targets=list(G), normalized=False)
        d2 = dl.copy()
        d2[s] = 0.0
        np.testing.assert_allclose(d2 * half, [ref[v] for v in range(n)], rtol=1e-12, atol=1e-12)


def test_bc_closed_forms(orc):
    # path 0-1-...-(n-1) from 0: one path to each vertex, delta[i] = n-1-i
    n = 40
    row, col = csr_of(GRAPHS["path"])
    lv, sg, dl = orc.bc_fp64(row, col, 0)
    np.testing.assert_array_equal(sg, np.ones(n))
    np.testing.assert_array_equal(dl, np.arange(n - 1, -1, -1, dtype=np.float64))
    # grid from a corner: sigma(x, y) = C(x + y, x)
    W, H = 13, 9
    row, col = csr_of(GRAPHS["grid"])
    _, sg, _ = orc.bc_fp64(row, col, 0)
    for x in range(W):
        for y in range(H):
            assert sg[x * H + y] == comb(x + y, x)
    # star: from a leaf, the hub carries every other leaf
    row, col = csr_of(GRAPHS["star"])
    _, _, dl = orc.bc_fp64(row, col, 1)
    assert dl[0] == 49.0 and np.all(dl[2:] == 0.0)
    # complete graph: no vertex lies inside a shortest path
    row, col = csr_of(GRAPHS["complete"])
    _, sg, dl = orc.bc_fp64(row, col, 3)
    assert np.all(dl[np.arange(12) != 3] == 0.0) and dl[3] == 11.0


# ---- PageRank ----------------------------------------------------------------------
def test_pr_regular_is_one(orc):
    for G in (nx.cycle_graph(30), nx.convert_node_labels_to_integers(nx.grid_2d_graph(6, 7, periodic=True))):
        row, col = csr_of(G)
        for variant in (0, 1):
            r, _ = orc.pr_fp64(row, col, variant, 0.85, 1e-12, 0)
            np.testing.assert_allclose(r, 1.0, atol=1e-10)


def test_pr_hand_iterated_path3(orc):
    row, col = csr_of(nx.path_graph(3))
    r, it = orc.pr_fp64(row, col, 0, 0.85, 0.0, -2)
    assert it == 2
    np.testing.assert_allclose(r, [0.21375, 0.405, 0.21375], rtol=1e-15)
    r, _ = orc.pr_fp64(row, col, 0, 0.85, 1e-13, 0)
    r0 = 0.21375 / 0.2775                              # r0 = .15 + .425 r1, r1 = .15 + 1.7 r0
    np.testing.assert_allclose(r, [r0, 0.15 + 1.7 * r0, r0], rtol=1e-11)


@pytest.mark.parametrize("name", ["ba", "grid", "complete"])
def test_pr_vs_networkx(orc, name):
    """without dangling vertices the unnormalised rule r = (1-d) + d sum r/deg is N times
    networkx's normalised PageRank"""
    G = GRAPHS[name]
    row, col = csr_of(G)
    n = len(row) - 1
    ref = nx.pagerank(G, alpha=0.85, tol=1e-14, max_iter=10000, weight=None)
    ref = np.array([ref[v] for v in range(n)]) * n
    for variant in (0, 1):
        r, _ = orc.pr_fp64(row, col, variant, 0.85, 1e-12, 0)
        np.testing.assert_allclose(r, ref, rtol=1e-8)


def test_pr_float_rule_tracks_double(orc):
    G = GRAPHS["ba"]
    row, col = csr_of(G)
    for variant in (0, 1):
        r32, it = orc.pr_f32(row, col, variant, 0.85, 0.01, 0)
        r64, _ = orc.pr_fp64(row, col, variant, 0.85, 0.01, -it)
        assert np.abs(r32 - r64).sum() / np.abs(r64).sum() < 1e-5


def test_pr_stopping_rule_invariant(orc):
    """PR_T exits when no vertex moves by more than tol: one more step moves none by more"""
    G = GRAPHS["ba"]
    row, col = csr_of(G)
    r, it = orc.pr_fp64(row, col, 0, 0.85, 0.01, 0)
    r1, _ = orc.pr_fp64(row, col, 0, 0.85, 0.0, -(it + 1))
    r0, _ = orc.pr_fp64(row, col, 0, 0.85, 0.0, -it)
    assert it > 3 and np.abs(r1 - r0).max() <= 0.011


# ---- planted mistakes: the checks above reject plausible slips -----------------------
def test_planted_mistakes_are_caught(orc):
    G = GRAPHS["directed"]
    row, col = csr_of(G)
    n = len(row) - 1
    # BC with the predecessor (transposed) direction, the form that is only right on
    # symmetric graphs, disagrees with networkx on a directed graph
    Gt = G.reverse()
    rt, ct = csr_of(Gt)
    s = 0
    ref = nx.betweenness_centrality_subset(G, sources=[s], targets=list(G), normalized=False)
    _, _, dl_wrong = orc.bc_fp64(rt, ct, s)
    dl_wrong[s] = 0
    assert not np.allclose(dl_wrong, [ref[v] for v in range(n)])
    # a damping slip (0.84) fails the networkx PageRank check
    Gb = GRAPHS["ba"]
    rb, cb = csr_of(Gb)
    nb = len(rb) - 1
    pr = nx.pagerank(Gb, alpha=0.85, tol=1e-14, max_iter=10000, weight=None)
    r_wrong, _ = orc.pr_fp64(rb, cb, 0, 0.84, 1e-12, 0)
    assert not np.allclose(r_wrong, np.array([pr[v] for v in range(nb)]) * nb, rtol=1e-8)
    # an off-by-one level fails the BFS check
    lv = orc.bfs(rb, cb, 0).copy()
    lv[5] += 1
    ref_l = nx.single_source_shortest_path_length(Gb, 0)
    assert not np.array_equal(lv, [ref_l[v] for v in range(nb)])


# ---- OpenMP drivers: bit-identical to the single-thread functions -----------------
@pytest.mark.parametrize("name", NAMES)
def test_multi_drivers_equal_singles(orc, name):
    G = GRAPHS[name]
    row, col = csr_of(G)
    w = weights_of(G, row, col)
    n = len(row) - 1
    src = np.array([0, n // 2, n - 1, n // 3, 0], np.int32)  # a repeated source too
    lv = orc.bfs_multi(row, col, src)
    dd = orc.sssp_dijkstra_multi(row, col, w, src)
    blv, bsg, bc = orc.bc_fp64_multi(row, col, src)
    ref_bc = np.zeros(n)
    for i, s in enumerate(src):
        s = int(s)
        assert np.array_equal(lv[i], orc.bfs(row, col, s))
        assert np.array_equal(dd[i], orc.sssp_dijkstra(row, col, w, s))
        l1, s1, d1 = orc.bc_fp64(row, col, s)
        assert np.array_equal(blv[i], l1) and np.array_equal(bsg[i], s1)
        d1 = d1.copy()
        d1[s] = 0.0
        ref_bc += d1
    assert np.array_equal(bc, ref_bc)  # same additions in the same (source) order


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("variant", [0, 1])
def test_pr_pull_equals_push(orc, name, variant):
    G = GRAPHS[name]
    row, col = csr_of(G)
    for tol, cap in ((1e-7, 100), (0.01, 0), (1e-9, -37)):
        r1, i1 = orc.pr_fp64(row, col, variant, 0.85, tol, cap)
        r2, i2 = orc.pr_fp64_pull(row, col, variant, 0.85, tol, cap)
        assert i1 == i2
        assert np.array_equal(r1, r2)


def test_omp_threads_reported(orc):
    assert orc.threads() >= 1


# ---- graphs past 2^31 - 1 edges: 64-bit row offsets ---------------------------------
@pytest.mark.parametrize("name", NAMES)
def test_bfs64_vs_networkx(orc, name):
    """orc_bfs64 (int64 row offsets) gives networkx's shortest-path lengths"""
    G = GRAPHS[name]
    row, col = csr_of(G)
    for s in _sources(G.number_of_nodes()):
        lv = orc.bfs(row.astype(np.int64), col, s)
        ref = np.full(G.number_of_nodes(), INF, np.int64)
        for v, d in nx.single_source_shortest_path_length(G, s).items():
            ref[v] = d
        np.testing.assert_array_equal(lv, ref)


@pytest.mark.parametrize("name", [n for n in NAMES if n != "directed"])
def test_sssp_certificate(orc, name):
    """The certificate holds for networkx's Dijkstra distances and fails for each kind
    of wrong answer: a raised distance, a lowered one, a reached vertex marked
    unreached, an unreached one given a distance, a nonzero source."""
    G = GRAPHS[name]
    row, col = csr_of(G)
    w = weights_of(G, row, col)
    n = G.number_of_nodes()
    for s in _sources(n):
        d = np.full(n, INF, np.int32)
        for v, x in nx.single_source_dijkstra_path_length(G, s, weight="w").items():
            d[v] = x
        lv = orc.bfs(row, col, s)
        assert orc.sssp_certificate(row, col, w, s, d, lv) == 0
        far = int(np.argmax(np.where(d != INF, d, -1)))        # a reached vertex other than s
        bad = d.copy(); bad[far] += 1
        assert orc.sssp_certificate(row, col, w, s, bad, lv) != 0
        bad = d.copy(); bad[far] -= 1
        assert orc.sssp_certificate(row, col, w, s, bad, lv) != 0
        bad = d.copy(); bad[far] = INF
        assert orc.sssp_certificate(row, col, w, s, bad, lv) != 0
        bad = d.copy(); bad[s] = 1
        assert orc.sssp_certificate(row, col, w, s, bad, lv) == 1
        un = np.flatnonzero(d == INF)
        if len(un):
            bad = d.copy(); bad[un[0]] = 5
            assert orc.sssp_certificate(row, col, w, s, bad, lv) != 0
