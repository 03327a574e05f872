"""Many small random inputs, degenerate ones included (no edges, one vertex, self-loops
only, duplicates, directed and symmetrized, several components): every engine and the
batched runs against the oracle on the same CSR.  Integer outputs bit-exact, BC
dependencies within 1e-6 of max(1, |delta|), PageRank within 1e-6 relative L1 of FP64
for the same number of iterations."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = 48


def _graph(mg, seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(1, 80))
    m = int(rng.integers(0, 4 * n + 1)) if seed % 7 else 0  # every 7th: no edges at all
    sym = seed % 3 != 0
    src = rng.integers(0, n, m).astype(np.int32)
    dst = rng.integers(0, n, m).astype(np.int32)
    if seed % 5 == 0:
        dst = src.copy()  # self-loops only
    val = rng.integers(1, 65, m).astype(np.int32)
    return mg.Graph.from_coo(n, src, dst, val, flags=mg.READ_SYM if sym else 0), sym, rng


@pytest.mark.parametrize("seed", range(CASES))
def test_random_small(mg, orc, seed):
    g, sym, rng = _graph(mg, seed)
    row, col, w = g.csr()
    nv = g.nv
    print(f"random seed={seed} nv={nv} ne={g.ne} sym={sym}", flush=True)
    if nv == 0:
        for algo in (mg.CC, mg.PR_T):
            e = mg.Engine(g, algo)
            e.run(0)
            assert len(e.result(mg.RES_RANK if algo == mg.PR_T else mg.RES_LEVEL)) == 0
            e.close()
        for algo in (mg.BFS, mg.SSSP, mg.BC):  # nothing to reach: the run ends at once
            e = mg.Engine(g, algo)
            e.run(0)
            assert len(e.result(mg.RES_LEVEL)) == 0
            e.close()
        return
    srcs = sorted({0, nv - 1, int(rng.integers(0, nv))})
    for algo, ref in ((mg.BFS, lambda s: orc.bfs(row, col, s)), (mg.SSSP, lambda s: orc.sssp_bf(row, col, w, s)[0])):
        e = mg.Engine(g, algo)
        for s in srcs:
            e.run(s)
            np.testing.assert_array_equal(e.result(mg.RES_LEVEL), ref(s), err_msg=f"algo {algo} src {s}")
        e.close()
    e = mg.Engine(g, mg.BC)
    for s in srcs:
        e.run(s)
        lv, sg, dl = orc.bc_fp64(row, col, s)
        np.testing.assert_array_equal(e.result(mg.RES_LEVEL), lv)
        np.testing.assert_array_equal(e.result(mg.RES_SIGMA), sg)
        np.testing.assert_allclose(e.result(mg.RES_DELTA).astype(np.float64), dl, rtol=1e-6, atol=1e-6)
    e.close()
    e = mg.Engine(g, mg.CC)
    e.run(0)
    np.testing.assert_array_equal(e.result(mg.RES_LEVEL), orc.cc_unionfind(row, col)[0])
    e.close()
    e = mg.Engine(g, mg.PR_D, pr_tolerance=1e-7, pr_max_iter=60)
    e.run(0)
    its = e.iterations()
    r = e.result(mg.RES_RANK).astype(np.float64)
    ref, _ = orc.pr_fp64(row, col, 1, 0.85, 1e-7, -its)
    assert np.abs(r - ref).sum() <= 1e-6 * max(np.abs(ref).sum(), 1e-30)
    e.close()
    # batches: up to 64 sources, repeats allowed
    k = int(rng.integers(1, 65))
    bs = rng.integers(0, nv, k).astype(np.int32)
    kinds = [mg.BATCH_BFS, mg.BATCH_BC] + ([mg.BATCH_SSSP] if sym else [])
    for kind in kinds:
        b = mg.Batch(g, kind)
        b.run(bs)
        lv = b.levels()
        for i in range(0, k, max(1, k // 6)):
            s = int(bs[i])
            want = orc.sssp_bf(row, col, w, s)[0] if kind == mg.BATCH_SSSP else orc.bfs(row, col, s)
            np.testing.assert_array_equal(lv[i], want, err_msg=f"batch kind {kind} row {i} src {s}")
        b.close()
