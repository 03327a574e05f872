"""The autotuner: it returns options from its candidate sets, never slower than the
starting point on its own measurement, and engines built with the tuned options
still match the oracle (tuning moves thresholds and shapes, never results)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_engine(mg, orc):
    g = mg.Graph.rmat(13, 16, seed=21)
    row, col, w = g.csr()
    src = [g.max_degree_vertex(), 5, 700]
    for algo in (mg.BFS, mg.SSSP, mg.BC):
        opt, t0, t1 = mg.autotune(g, algo, src, reps=1)
        assert t1 <= t0 + 1e-6
        assert opt.alpha in (0, 4, 8, 15, 30, 60) and opt.beta in (0, 6, 12, 18, 36, 72)
        assert opt.grid_blocks in (0, 1024, 2048, 4096) and opt.graph_levels in (0, 4, 16)
        # bin edges and block shape (0: the default, kept when no candidate was faster)
        assert opt.huge_degree in (4096, 8192, 16384, 32768) and opt.block_threads in (0, 256, 512, 1024)
        if algo != mg.SSSP:
            assert opt.pull_split in (0, 1024, 2048, 4096, 8192)
        if algo == mg.BC:
            assert opt.lane_edges in (0, 16, 32, 64) and opt.back_split in (0, 1024, 2048, 4096)
            assert opt.back_chunk in (0, 512, 1024, 2048)
        print(f"algo {algo}: {t0:.3f} -> {t1:.3f} ms, block {opt.block_threads}, huge {opt.huge_degree}, "
              f"pull_split {opt.pull_split}, lane {opt.lane_edges}, back {opt.back_split}/{opt.back_chunk}", flush=True)
        e = mg.Engine(g, algo, options=opt)
        for s in src:
            e.run(s)
            if algo == mg.BC:
                lv, sg, dl = orc.bc_fp64(row, col, s)
                np.testing.assert_array_equal(e.result(mg.RES_SIGMA), sg)
                np.testing.assert_allclose(e.result(mg.RES_DELTA).astype(np.float64), dl, rtol=1e-6, atol=1e-6)
                continue
            ref = orc.bfs(row, col, s) if algo == mg.BFS else orc.sssp_bf(row, col, w, s)[0]
            np.testing.assert_array_equal(e.result(mg.RES_LEVEL), ref)
        e.close()


def test_autotune_pagerank(mg, orc):
    g = mg.Graph.rmat(14, 8, seed=22)
    row, col, _ = g.csr()
    opt, t0, t1 = mg.autotune(g, mg.PR_D, [0], reps=1, pr_tolerance=1e-7, pr_max_iter=40)
    assert t1 <= t0 + 1e-6 and opt.part_log2 in (0, 11, 12, 13)
    e = mg.Engine(g, mg.PR_D, options=opt)
    e.run(0)
    ref, _ = orc.pr_fp64(row, col, 1, 0.85, 1e-7, -e.iterations())
    r = e.result(mg.RES_RANK).astype(np.float64)
    assert np.abs(r - ref).sum() / np.abs(ref).sum() <= 1e-6
    e.close()


def test_autotune_batch_union(mg, orc):
    parts = [mg.Graph.rmat(11, 8, seed=23), mg.Graph.grid(60, 50), mg.Graph.rmat(9, 16, seed=24)]
    u = mg.Graph.union(parts)
    rng = np.random.default_rng(25)
    src, slots, offs, off = [], [], [], 0
    for p in parts:
        offs.append(off)
        src += (rng.integers(0, p.nv, 4) + off).tolist()
        slots += [0, 1, 2, 3]
        off += p.nv
    for kind in (mg.BATCH_BFS, mg.BATCH_SSSP):
        opt, t0, t1 = mg.batch_autotune(u, kind, src, slots, reps=1)
        assert t1 <= t0 + 1e-6 and opt.batch_chunk in (0, 64, 128, 256, 512, 1024, 4096)
        if kind == mg.BATCH_BFS:
            assert opt.batch_light in (0, 16, 32, 64)
        assert opt.sparse_grid in (0, 64, 128)
        b = mg.Batch(u, kind, options=opt)
        b.run(src, slots)
        lv = b.levels()
        for gi, p in enumerate(parts):
            row, col, w = p.csr()
            o = offs[gi]
            for j in range(4):
                s = src[4 * gi + j] - o
                ref = orc.bfs(row, col, s) if kind == mg.BATCH_BFS else orc.sssp_bf(row, col, w, s)[0]
                np.testing.assert_array_equal(lv[j, o:o + p.nv], ref)
        b.close()
