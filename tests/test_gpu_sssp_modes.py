"""SSSP through every level mix: push only, pull only, and the switching hybrid,
eager and graph-captured, against Bellman-Ford (bit-exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODES = [dict(), dict(mode=1), dict(mode=2), dict(graph_levels=3), dict(mode=2, graph_levels=2), dict(alpha=1000.0)]


@pytest.mark.parametrize("opts", MODES, ids=["hybrid", "push", "pull", "graph", "pull-graph", "eager-pull-switch"])
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
