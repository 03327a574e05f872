"""The C ABI on the CPU: libmgamd.so loads, exports every function that
include/multigraph_amd.h declares, and its host-side error paths return the
documented status codes.  No compute call reaches a GPU here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "multigraph_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(mg_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("mg_graph_load_mtx", "mg_build", "mg_reset", "mg_step", "mg_bc_backward", "mg_run",
                 "mg_get_result", "mg_release", "mg_batch_build", "mg_batch_run", "mg_batch_get"):
        assert must in names
    assert len(names) >= 30


def test_every_declared_symbol_is_exported(mg):
    L = C.CDLL(mg.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_version_and_defaults(mg):
    L = mg.lib()
    L.mg_version.restype = C.c_char_p
    assert L.mg_version().decode().startswith("multigraph-amd")
    o = mg.default_options()
    assert o.mode == mg.MODE_HYBRID and o.alpha == 0.0 and o.beta == 0.0 and o.grid_blocks == 0  # the paths' own
    assert abs(o.pr_damping - 0.85) < 1e-7 and abs(o.pr_tolerance - 0.01) < 1e-9   # common.h:59-60
    assert o.pr_max_iter == 0 and o.graph_levels == 0


def test_null_and_range_errors(mg):
    L = mg.lib()
    h = C.c_void_p()
    assert L.mg_graph_load_mtx(None, 0, 1, C.c_uint64(1), C.byref(h)) == mg.ERR_ARG
    assert L.mg_build(None, mg.BFS, None, C.byref(h)) == mg.ERR_ARG
    assert L.mg_release(None) == mg.ERR_ARG
    assert L.mg_batch_release(None) == mg.ERR_ARG
    L.mg_last_error.restype = C.c_char_p
    assert L.mg_last_error()
    g = mg.Graph.rmat(6, 4, seed=1)
    # an unknown algorithm or batch kind is refused before any device work
    assert L.mg_build(g._h, 17, None, C.byref(h)) == mg.ERR_ARG
    assert L.mg_batch_build(g._h, 9, 64, None, C.byref(h)) == mg.ERR_ARG
    assert L.mg_batch_build(g._h, mg.BATCH_BFS, 65, None, C.byref(h)) == mg.ERR_LIMIT
    with pytest.raises(mg.MGError):
        g.remap(1 << 20)


def test_device_count_without_gpu(mg):
    L = mg.lib()
    n = C.c_int(-1)
    rc = L.mg_device_count(C.byref(n))
    assert rc in (mg.OK, mg.ERR_DEVICE)
    if rc == mg.OK:
        assert n.value >= 0


def test_graph_info_and_csr_views(mg):
    g = mg.Graph.grid(4, 3)
    nv, ne, sp = g.info()
    assert (nv, ne, sp) == (12, 34, 0)
    row, col, w = g.csr()
    assert row.dtype == np.int32 and len(row) == nv + 1 and len(col) == ne == len(w)
    np.testing.assert_array_equal(g.orig_ids(), np.arange(12))


def test_graph_from_csr(mg):
    """a caller-built CSR: ids kept, rows sorted by column (weights carried), a repeated
    column keeps its first entry; bad offsets and columns are refused"""
    row = np.array([0, 3, 5, 7, 8])
    col = np.array([2, 1, 2, 0, 3, 1, 0, 1])
    w = np.array([7, 5, 9, 5, 4, 4, 7, 6])
    g = mg.Graph.from_csr(row, col, w)
    r, c, ww = g.csr()
    assert g.nv == 4
    np.testing.assert_array_equal(r, [0, 2, 4, 6, 7])
    np.testing.assert_array_equal(c, [1, 2, 0, 3, 0, 1, 1])
    np.testing.assert_array_equal(ww, [5, 7, 5, 4, 7, 4, 6])   # row 0: column 2 keeps its first weight, 7
    g2 = mg.Graph.from_csr([0, 1, 1], [1])                      # unweighted: weight 1
    np.testing.assert_array_equal(g2.csr()[2], [1])
    with pytest.raises(mg.MGError) as e:
        mg.Graph.from_csr([0, 2, 1], [1, 0])
    assert e.value.code == mg.ERR_ARG
    with pytest.raises(mg.MGError) as e:
        mg.Graph.from_csr([0, 1, 2], [1, 5])
    assert e.value.code == mg.ERR_RANGE


def test_graph_from_csr_symmetry(mg):
    """the symmetric property is detected, weights included: an edge whose reverse has
    another weight makes the graph directed (SSSP pulls read a row's weights as the
    in-edges')"""
    assert mg.Graph.from_csr([0, 1, 3, 4], [1, 0, 2, 1]).symmetric
    assert not mg.Graph.from_csr([0, 1, 2, 2], [1, 2]).symmetric
    assert mg.Graph.from_csr([0, 1, 3, 4], [1, 0, 2, 1], [5, 5, 7, 7]).symmetric
    assert not mg.Graph.from_csr([0, 1, 3, 4], [1, 0, 2, 1], [5, 5, 7, 8]).symmetric
    assert mg.Graph.rmat(8, 4, seed=2).symmetric
    assert not mg.Graph.rmat(8, 4, seed=2, flags=0).symmetric
