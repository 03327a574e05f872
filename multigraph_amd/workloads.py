"""The benchmark workloads, built the same way by bench.py and the full-size parity tests.

rmat(mg, scale)       symmetrized RMAT (a=.57 b=.19 c=.19, edge factor 16, seed 1), edge
                      weights uniform int in [1, 64] from seed 7; self-loops and duplicates
                      dropped, isolated vertices compacted (the reader's rules)
batch_sources(nv)     64 sources uniform over the compacted vertices (numpy default_rng(3))
uneven(mg)            32 graphs joined as one disjoint union: RMAT-14..20 with 4 seeds each
                      and 4-neighbour grids 256^2, 512^2, 1024^2, 2048^2; 4 sources per
                      graph (default_rng(5)) sharing slots 0..3 of one batched run
"""
from __future__ import annotations

import numpy as np

RMAT_ABC = (0.57, 0.19, 0.19)
UNEVEN_DESC = "RMAT-14..20 x 4 seeds + grids 256^2, 512^2, 1024^2, 2048^2"


def rmat(mg, scale: int, edge_factor: int = 16):
    a, b, c = RMAT_ABC
    return mg.Graph.rmat(scale, edge_factor, a, b, c, seed=1, weight_seed=7)


def batch_sources(nv: int, n: int = 64, seed: int = 3) -> np.ndarray:
    return np.random.default_rng(seed).integers(0, nv, size=n).astype(np.int32)


def uneven_members(mg):
    graphs = [mg.Graph.rmat(sc, 16, seed=100 + 10 * k + sc, weight_seed=7) for k in range(4) for sc in range(14, 21)]
    graphs += [mg.Graph.grid(side, side, weight_seed=7) for side in (256, 512, 1024, 2048)]
    return graphs


def uneven(mg):
    """(union graph, member graphs, member offsets, sources, slots)"""
    graphs = uneven_members(mg)
    u = mg.Graph.union(graphs)
    rng = np.random.default_rng(5)
    src, slots, offs, off = [], [], [], 0
    for g in graphs:
        offs.append(off)
        src += (rng.integers(0, g.nv, 4) + off).tolist()
        slots += [0, 1, 2, 3]
        off += g.nv
    return u, graphs, offs, np.asarray(src, np.int32), np.asarray(slots, np.int32)
