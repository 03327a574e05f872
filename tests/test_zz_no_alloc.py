"""The allocation-free BFS entry point (mg_bfs_noalloc) against the engine's BFS
(mg_build + mg_run), through the C ABI, on two calls in a row with the second on a
smaller graph, plus the oracle.  Device buffers come from the HIP runtime the library
itself is linked with."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(int(nbytes), 4)) == 0
        self.nbytes = int(nbytes)

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if a.nbytes:
            assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, n, dtype):
        out = np.empty(n, dtype)
        if out.nbytes:
            assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _run_noalloc(mg, hip, row, col, src, mode=0):
    nv, ne = len(row) - 1, len(col)
    dr, dc, dl = DevBuf(hip, 4 * (nv + 1)), DevBuf(hip, 4 * ne), DevBuf(hip, 4 * nv)
    try:
        dr.put(row.astype(np.int32))
        dc.put(col.astype(np.int32))
        dl.put(np.full(nv, 7, np.int32))  # stale contents must be overwritten
        mg.bfs_noalloc(nv, ne, dr.p.value, dc.p.value, src, dl.p.value, mode)
        return dl.get(nv, np.int32)
    finally:
        for b in (dr, dc, dl):
            b.free()


def _engine_bfs(mg, g, src, **kw):
    e = mg.Engine(g, mg.BFS, **kw)
    e.run(src)
    lv = e.result(mg.RES_LEVEL)
    e.close()
    return lv


def test_noalloc_matches_engine_two_calls(mg, orc):
    hip = _hip()
    print("noalloc: RMAT-18 hub, then RMAT-12 vertex 5", flush=True)
    cases = [(mg.Graph.rmat(18, 16, seed=1), None), (mg.Graph.rmat(12, 8, seed=4), 5)]
    for g, src in cases:
        row, col, _ = g.csr()
        s = g.max_degree_vertex() if src is None else src
        got = _run_noalloc(mg, hip, row, col, s)
        ref = _engine_bfs(mg, g, s)
        assert np.array_equal(got, ref), f"noalloc levels differ from mg_run on nv={g.nv}"
        assert np.array_equal(got, orc.bfs(row, col, s))


def test_noalloc_grid_and_directed(mg, orc):
    hip = _hip()
    print("noalloc: 300x200 grid corner; directed RMAT-14 push-only", flush=True)
    g = mg.Graph.grid(300, 200)
    row, col, _ = g.csr()
    got = _run_noalloc(mg, hip, row, col, 0)
    assert np.array_equal(got, _engine_bfs(mg, g, 0))
    assert got.max() == 299 + 199  # Manhattan distance of the far corner
    gd = mg.Graph.rmat(14, 8, seed=9, flags=0)
    row, col, _ = gd.csr()
    s = gd.max_degree_vertex()
    got = _run_noalloc(mg, hip, row, col, s, mode=mg.MODE_SPARSE)
    assert np.array_equal(got, orc.bfs(row, col, s))
    assert np.array_equal(got, _engine_bfs(mg, gd, s))


def test_noalloc_rejects_oversize(mg):
    L = mg.lib()
    assert L.mg_bfs_noalloc((1 << 22) + 1, 10, None, None, 0, 0, None, None) == mg.ERR_LIMIT
    assert L.mg_bfs_noalloc(10, 1 << 31, None, None, 0, 0, None, None) == mg.ERR_LIMIT
    assert L.mg_bfs_noalloc(10, 10, None, None, 0, 0, None, None) == mg.ERR_ARG
