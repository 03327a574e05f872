"""SSSP through every level mix: push only, pull only, and the switching hybrid,
eager and graph-captured, against Bellman-Ford (bit-exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODES = [dict(), dict(mode=1), dict(mode=2), dict(graph_levels=3), dict(mode=2, graph_levels=2), dict(alpha=1000.0),
         dict(pull_layout=1), dict(mode=2, pull_layout=1), dict(mode=2, pull_layout=1, graph_levels=4), dict(part_log2=6),
         dict(mode=2, part_log2=5), dict(pull_layout=2), dict(mode=2, pull_layout=2)]
IDS = ["hybrid", "push", "pull", "graph", "pull-graph", "eager-pull-switch", "csr", "csr-pull", "csr-pull-graph",
       "small-parts", "pull-tiny-parts", "wide-msgs", "wide-msgs-pull"]


@pytest.mark.parametrize("opts", MODES, ids=IDS)
@pytest.mark.parametrize("gname", ["rmat14", "grid", "rmat12x2"])
def test_sssp_modes(mg, orc, opts, gname):
    g = {"rmat14": lambda: mg.Graph.rmat(14, 16, seed=3),
         "grid": lambda: mg.Graph.grid(40, 70),
         "rmat12x2": lambda: mg.Graph.rmat(12, 2, seed=9)}[gname]()
    row, col, w = g.csr()
    e = mg.Engine(g, mg.SSSP, **opts)
    for s in (0, g.max_degree_vertex(), g.nv - 1):
        print(f"sssp {gname} {opts} src={s}", flush=True)
        e.run(s)
        ref, _ = orc.sssp_bf(row, col, w, s)
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), ref)
    e.close()


@pytest.mark.parametrize("opts", [dict(mode=2), dict(mode=2, part_log2=7), dict(mode=2, pull_layout=1), dict()],
                         ids=["layout", "layout-p128", "csr", "hybrid"])
def test_sssp_pull_union(mg, orc, opts):
    """dense sweeps on a disjoint union (hub rows split over several gather units at
    small partitions, members with and without edges in a partition): distances
    bit-exact, through the partitioned layout and the CSR"""
    parts = [mg.Graph.rmat(12, 16, seed=41), mg.Graph.grid(33, 29), mg.Graph.rmat(10, 4, seed=42)]
    u = mg.Graph.union(parts)
    row, col, w = u.csr()
    e = mg.Engine(u, mg.SSSP, **opts)
    for s in (0, parts[0].nv + 5, u.nv - 1):
        e.run(s)
        ref, _ = orc.sssp_bf(row, col, w, s)
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), ref, err_msg=f"{opts} source {s}")
    e.close()


@pytest.mark.parametrize("mode", [0, 2])
def test_sssp_layout_split_partition(mg, orc, mode):
    """a hub with 50 K neighbours: its destination partition's edge stream is cut over
    several gather units (partial minima lowered with atomics, messages rebuilt by a
    separate pass); distances bit-exact from the hub, a leaf and a far vertex"""
    rng = np.random.default_rng(77)
    n = 60000
    src = [0] * 50000 + rng.integers(1, n, 40000).tolist()
    dst = list(range(1, 50001)) + rng.integers(1, n, 40000).tolist()
    val = rng.integers(1, 65, len(src)).tolist()
    g = mg.Graph.from_coo(n, np.array(src), np.array(dst), np.array(val))
    row, col, w = g.csr()
    e = mg.Engine(g, mg.SSSP, mode=mode)
    for s in (0, 17, g.nv - 1):
        e.run(s)
        ref, _ = orc.sssp_bf(row, col, w, s)
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), ref, err_msg=f"mode {mode} source {s}")
    e.close()


@pytest.mark.parametrize("wmax", [255, 60000, 100000], ids=["w8", "w16", "csr-fallback"])
def test_sssp_weight_widths(mg, orc, wmax):
    """the layout stores weights as bytes up to 255, as 16 bits up to 65535, and larger
    weights fall back to the CSR sweeps: distances bit-exact in every case"""
    rng = np.random.default_rng(wmax)
    n = 3000
    src = rng.integers(0, n, 24000)
    dst = rng.integers(0, n, 24000)
    val = rng.integers(1, wmax + 1, 24000)
    g = mg.Graph.from_coo(n, src, dst, val)
    row, col, w = g.csr()
    assert w.max() <= wmax
    e = mg.Engine(g, mg.SSSP, mode=2)
    for s in (0, g.nv // 2):
        e.run(s)
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), orc.sssp_bf(row, col, w, s)[0], err_msg=f"wmax {wmax} src {s}")
    e.close()


@pytest.mark.parametrize("opts", [dict(), dict(mode=2), dict(mode=2, graph_levels=3), dict(mode=2, pull_layout=2),
                                  dict(mode=2, part_log2=6)],
                         ids=["hybrid", "pull", "pull-graph", "pull-wide", "pull-small-parts"])
def test_sssp_message_widths(mg, orc, opts):
    """layout sweeps send 16-bit distance messages until a finite distance above 65534
    must be sent: that vertex is stamped (it stays in the frontier) and the run goes
    on with 32-bit messages; captured sweeps and pull_layout=2 use 32-bit throughout.
    Weights up to 30000 push distances past 65534 within a few hops; a 1,500-vertex
    path with weight-64 edges does it after ~1,000 hops, mid-run."""
    rng = np.random.default_rng(41)
    n = 2000
    src = rng.integers(0, n, 16000)
    dst = rng.integers(0, n, 16000)
    val = rng.integers(1, 30001, 16000)
    k = 1500
    src = np.concatenate([src, np.arange(n, n + k - 1)])
    dst = np.concatenate([dst, np.arange(n + 1, n + k)])
    val = np.concatenate([val, np.full(k - 1, 64)])
    src = np.concatenate([src, [0]])
    dst = np.concatenate([dst, [n]])
    val = np.concatenate([val, [5]])
    g = mg.Graph.from_coo(n + k, src, dst, val)
    row, col, w = g.csr()
    e = mg.Engine(g, mg.SSSP, **opts)
    for s in (0, g.nv - 1, 7):
        print(f"sssp widths {opts} src={s}", flush=True)
        e.run(s)
        ref, _ = orc.sssp_bf(row, col, w, s)
        assert ref[ref != 0x64646464].max() > 65535
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), ref, err_msg=f"{opts} source {s}")
    e.close()


def test_profile_filter(mg, orc):
    """per-kernel accounting: every kernel by default, only the named prefix with a
    filter; the run's result is the same either way"""
    g = mg.Graph.rmat(14, 16, seed=3)
    row, col, w = g.csr()
    s = g.max_degree_vertex()
    e = mg.Engine(g, mg.SSSP, mode=2)  # every level a layout sweep
    e.set_profiling(True)
    e.run(s)
    full = e.kernel_stats()
    assert "k_ps_gather" in full and len(full) > 1
    e.set_profiling(True, only="k_ps_gather")
    e.run(s)
    only = e.kernel_stats()
    assert set(only) == {"k_ps_gather"} and only["k_ps_gather"]["launches"] == full["k_ps_gather"]["launches"]
    np.testing.assert_array_equal(e.result(mg.RES_LEVEL), orc.sssp_bf(row, col, w, s)[0])
    e.set_profiling(False)
    e.close()
