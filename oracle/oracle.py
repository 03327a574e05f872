"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
INF = 0x64646464
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        I = C.POINTER(C.c_int32)
        L.orc_bfs.argtypes = [C.c_int, I, I, C.c_int, I]
        L.orc_sssp_bf.argtypes = [C.c_int, I, I, I, C.c_int, I]
        L.orc_sssp_dijkstra.argtypes = [C.c_int, I, I, I, C.c_int, I]
        L.orc_canon.argtypes = [C.c_int, I, I]
        L.orc_cc_labelprop.argtypes = [C.c_int, I, I, I, C.POINTER(C.c_int)]
        L.orc_cc_unionfind.argtypes = [C.c_int, I, I, I]
        D = C.POINTER(C.c_double)
        F = C.POINTER(C.c_float)
        L.orc_bc_fp64.argtypes = [C.c_int, I, I, C.c_int, I, D, D]
        L.orc_pr_f32.argtypes = [C.c_int, I, I, C.c_int, C.c_float, C.c_float, C.c_int, F]
        L.orc_pr_fp64.argtypes = [C.c_int, I, I, C.c_int, C.c_double, C.c_double, C.c_int, D]
        L.orc_threads.argtypes = []
        L.orc_bfs_multi.argtypes = [C.c_int, I, I, C.c_int, I, I]
        L.orc_sssp_dijkstra_multi.argtypes = [C.c_int, I, I, I, C.c_int, I, I]
        L.orc_bc_fp64_multi.argtypes = [C.c_int, I, I, C.c_int, I, I, D, D]
        L.orc_pr_fp64_pull.argtypes = [C.c_int, I, I, C.c_int, C.c_double, C.c_double, C.c_int, D]
        I64 = C.POINTER(C.c_int64)
        L.orc_bfs64.argtypes = [C.c_int, I64, I, C.c_int, I]
        L.orc_sssp_certificate64.argtypes = [C.c_int, I64, I, I, C.c_int, I, I]
        _lib = L
    return _lib


def _p(a, t=C.c_int32):
    return a.ctypes.data_as(C.POINTER(t))


def _csr(row, col):
    return np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32)


def bfs(row, col, src):
    """BFS levels; 64-bit row offsets (a graph past 2^31 - 1 edges) go to orc_bfs64"""
    nv = len(row) - 1
    out = np.empty(nv, np.int32)
    if np.asarray(row).dtype == np.int64:
        row = np.ascontiguousarray(row, np.int64)
        col = np.ascontiguousarray(col, np.int32)
        lib().orc_bfs64(nv, _p(row, C.c_int64), _p(col), src, _p(out))
        return out
    row, col = _csr(row, col)
    lib().orc_bfs(nv, _p(row), _p(col), src, _p(out))
    return out


def sssp_certificate(row, col, w, src, dist, level):
    """0 when `dist` are the exact shortest distances from src of a symmetric CSR
    (orc_sssp_certificate64: source 0, no edge lowers a distance, every reached vertex
    tight, reached set = the BFS levels' `level`); else the failed condition 1..4"""
    row = np.ascontiguousarray(row, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    w = np.ascontiguousarray(w, np.int32)
    dist = np.ascontiguousarray(dist, np.int32)
    level = np.ascontiguousarray(level, np.int32)
    return lib().orc_sssp_certificate64(len(row) - 1, _p(row, C.c_int64), _p(col), _p(w), src, _p(dist), _p(level))


def sssp_bf(row, col, w, src):
    row, col = _csr(row, col)
    w = np.ascontiguousarray(w, np.int32)
    nv = len(row) - 1
    out = np.empty(nv, np.int32)
    sweeps = lib().orc_sssp_bf(nv, _p(row), _p(col), _p(w), src, _p(out))
    return out, sweeps


def sssp_dijkstra(row, col, w, src):
    row, col = _csr(row, col)
    w = np.ascontiguousarray(w, np.int32)
    nv = len(row) - 1
    out = np.empty(nv, np.int32)
    lib().orc_sssp_dijkstra(nv, _p(row), _p(col), _p(w), src, _p(out))
    return out


def canon(labels):
    labels = np.ascontiguousarray(labels, np.int32)
    out = np.empty_like(labels)
    k = lib().orc_canon(len(labels), _p(labels), _p(out))
    if k < 0:
        raise ValueError("labels out of range")
    return out


def cc_labelprop(row, col):
    row, col = _csr(row, col)
    nv = len(row) - 1
    out = np.empty(nv, np.int32)
    sweeps = C.c_int(0)
    k = lib().orc_cc_labelprop(nv, _p(row), _p(col), _p(out), C.byref(sweeps))
    return out, k, sweeps.value


def cc_unionfind(row, col):
    row, col = _csr(row, col)
    nv = len(row) - 1
    out = np.empty(nv, np.int32)
    k = lib().orc_cc_unionfind(nv, _p(row), _p(col), _p(out))
    return out, k


def bc_fp64(row, col, src):
    row, col = _csr(row, col)
    nv = len(row) - 1
    lv = np.empty(nv, np.int32)
    sg = np.empty(nv, np.float64)
    dl = np.empty(nv, np.float64)
    lib().orc_bc_fp64(nv, _p(row), _p(col), src, _p(lv), _p(sg, C.c_double), _p(dl, C.c_double))
    return lv, sg, dl


def pr_f32(row, col, variant, damping=0.85, tol=0.01, max_iter=0):
    row, col = _csr(row, col)
    nv = len(row) - 1
    r = np.empty(nv, np.float32)
    it = lib().orc_pr_f32(nv, _p(row), _p(col), variant, damping, tol, max_iter, _p(r, C.c_float))
    return r, it


def pr_fp64(row, col, variant, damping=0.85, tol=0.01, max_iter=0):
    row, col = _csr(row, col)
    nv = len(row) - 1
    r = np.empty(nv, np.float64)
    it = lib().orc_pr_fp64(nv, _p(row), _p(col), variant, damping, tol, max_iter, _p(r, C.c_double))
    return r, it


# ---- OpenMP drivers (bit-identical to the single-thread functions above) ----------
def threads():
    """OpenMP threads the multi-source and pull functions use"""
    return lib().orc_threads()


def _srcs(sources):
    return np.ascontiguousarray(sources, np.int32)


def bfs_multi(row, col, sources):
    """[n][nv] levels: row s = bfs(row, col, sources[s])"""
    row, col = _csr(row, col)
    src = _srcs(sources)
    nv = len(row) - 1
    out = np.empty((len(src), nv), np.int32)
    if lib().orc_bfs_multi(nv, _p(row), _p(col), len(src), _p(src), _p(out)) < 0:
        raise MemoryError("orc_bfs_multi")
    return out


def sssp_dijkstra_multi(row, col, w, sources):
    """[n][nv] distances: row s = sssp_dijkstra(row, col, w, sources[s])"""
    row, col = _csr(row, col)
    w = np.ascontiguousarray(w, np.int32)
    src = _srcs(sources)
    nv = len(row) - 1
    out = np.empty((len(src), nv), np.int32)
    if lib().orc_sssp_dijkstra_multi(nv, _p(row), _p(col), _p(w), len(src), _p(src), _p(out)) < 0:
        raise MemoryError("orc_sssp_dijkstra_multi")
    return out


def bc_fp64_multi(row, col, sources):
    """(levels [n][nv], sigma [n][nv], bc [nv]): rows as bc_fp64 per source; bc sums the
    sources' dependencies, each source's own entry left out"""
    row, col = _csr(row, col)
    src = _srcs(sources)
    nv = len(row) - 1
    lv = np.empty((len(src), nv), np.int32)
    sg = np.empty((len(src), nv), np.float64)
    bc = np.empty(nv, np.float64)
    if lib().orc_bc_fp64_multi(nv, _p(row), _p(col), len(src), _p(src), _p(lv), _p(sg, C.c_double),
                               _p(bc, C.c_double)) < 0:
        raise MemoryError("orc_bc_fp64_multi")
    return lv, sg, bc


def pr_fp64_pull(row, col, variant, damping=0.85, tol=0.01, max_iter=0):
    """pr_fp64 in pull form on OpenMP threads: bit-identical ranks"""
    row, col = _csr(row, col)
    nv = len(row) - 1
    r = np.empty(nv, np.float64)
    it = lib().orc_pr_fp64_pull(nv, _p(row), _p(col), variant, damping, tol, max_iter, _p(r, C.c_double))
    if it < 0:
        raise MemoryError("orc_pr_fp64_pull")
    return r, it
