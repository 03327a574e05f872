"""SSSP through every level mix: push only, pull only, and the switching hybrid,
eager and graph-captured, against Bellman-Ford (bit-exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODES = [dict(), dict(mode=1), dict(mode=2), dict(graph_levels=3), dict(mode=2, graph_levels=2), dict(alpha=1000.0),
         dict(pull_groups=1), dict(pull_groups=3), dict(mode=2, pull_groups=7), dict(mode=2, pull_groups=64, graph_levels=4)]
IDS = ["hybrid", "push", "pull", "graph", "pull-graph", "eager-pull-switch", "one-group", "groups3", "pull-groups7",
       "pull-groups64-graph"]


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


@pytest.mark.parametrize("groups", [2, 5, 13])
def test_sssp_source_groups_union(mg, orc, groups):
    """pull sweeps over source groups on a disjoint union (some groups hold rows of
    one member only, group edge ranges start 8-aligned, a chunk can hold one row's
    edges from several chunks' worth of a hub): distances bit-exact"""
    parts = [mg.Graph.rmat(12, 16, seed=41), mg.Graph.grid(33, 29), mg.Graph.rmat(10, 4, seed=42)]
    u = mg.Graph.union(parts)
    row, col, w = u.csr()
    e = mg.Engine(u, mg.SSSP, mode=2, pull_groups=groups)
    for s in (0, parts[0].nv + 5, u.nv - 1):
        e.run(s)
        ref, _ = orc.sssp_bf(row, col, w, s)
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), ref, err_msg=f"groups {groups} source {s}")
    e.close()
