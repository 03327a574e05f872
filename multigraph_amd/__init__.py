"""multigraph_amd — MI355X (gfx950) frontier-processing engine, Python binding.

A thin ctypes layer over the C ABI in include/multigraph_amd.h (libmgamd.so, built
in-tree by `make -C multigraph_amd`).  There is no fallback: if the library is
missing, importing works but every call raises, and engine construction raises when
no HIP device is present.

    g = Graph.rmat(scale=16, edge_factor=16, seed=1)
    e = Engine(g, BFS)
    ms = e.run(g.start_p)
    levels = e.result(RES_LEVEL)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MG_LIBRARY selects another build of the same ABI (e.g. one compiled with a
# maintainer's own hooks file); the default is the in-tree libmgamd.so
LIB_PATH = os.environ.get("MG_LIBRARY") or os.path.join(_HERE, "libmgamd.so")

BFS, SSSP, BC, CC, PR_T, PR_D = range(6)
ALGO_NAMES = {BFS: "BFS", SSSP: "SSSP", BC: "BC", CC: "CC", PR_T: "PR_T", PR_D: "PR_D"}
READ_SYM, READ_CC, READ_HEADER = 1, 2, 4
MODE_HYBRID, MODE_SPARSE, MODE_DENSE = 0, 1, 2
RES_LEVEL, RES_SIGMA, RES_DELTA, RES_RANK = 0, 1, 2, 3
BATCH_BFS, BATCH_BC, BATCH_SSSP = 0, 1, 2
INF = 0x64646464
OK, ERR_ARG, ERR_IO, ERR_RANGE, ERR_NOMEM, ERR_DEVICE, ERR_LIMIT, ERR_STATE = 0, -1, -2, -3, -4, -5, -6, -7


class MGError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"multigraph_amd error {code}: {msg}")
        self.code = code


class Options(C.Structure):
    _fields_ = [
        ("mode", C.c_int),
        ("alpha", C.c_float),
        ("beta", C.c_float),
        ("huge_degree", C.c_int),
        ("grid_blocks", C.c_int),
        ("graph_levels", C.c_int),
        ("pr_damping", C.c_float),
        ("pr_tolerance", C.c_float),
        ("pr_max_iter", C.c_int),
        ("part_log2", C.c_int),
        ("batch_chunk", C.c_int),
        ("persist_levels", C.c_int),
        ("pull_layout", C.c_int),
        ("sssp_delta", C.c_int),
        ("batch_lanes", C.c_int),
        ("lane_edges", C.c_int),
        ("pull_split", C.c_int),
        ("back_split", C.c_int),
        ("back_chunk", C.c_int),
        ("block_threads", C.c_int),
        ("batch_light", C.c_int),
        ("sparse_grid", C.c_int),
        ("wide_offsets", C.c_int),
    ]


_lib = None

_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_SIGS = {
    "mg_last_error": (C.c_char_p, []),
    "mg_version": (C.c_char_p, []),
    "mg_graph_load_mtx": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_uint64, C.POINTER(_P)]),
    "mg_graph_from_csr": (C.c_int, [C.c_int, C.POINTER(C.c_int64), _I32P, _I32P, C.POINTER(_P)]),
    "mg_graph_from_coo": (C.c_int, [C.c_int, C.c_int64, _I32P, _I32P, _I32P, C.c_int, C.c_int, C.POINTER(_P)]),
    "mg_graph_gen_rmat": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_uint64,
                                    C.c_int, C.POINTER(_P)]),
    "mg_graph_gen_grid": (C.c_int, [C.c_int, C.c_int, C.c_uint64, C.c_int, C.POINTER(_P)]),
    "mg_graph_union": (C.c_int, [C.c_int, C.POINTER(_P), C.POINTER(_P)]),
    "mg_graph_write_mtx": (C.c_int, [_P, C.c_char_p]),
    "mg_graph_info": (C.c_int, [_P, _I32P, C.POINTER(C.c_int64), _I32P]),
    "mg_graph_symmetric": (C.c_int, [_P, _I32P]),
    "mg_graph_csr": (C.c_int, [_P, C.POINTER(_I32P), C.POINTER(_I32P), C.POINTER(_I32P)]),
    "mg_graph_csr64": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "mg_graph_orig_ids": (C.c_int, [_P, C.POINTER(_I32P)]),
    "mg_graph_remap": (C.c_int, [_P, C.c_int, _I32P]),
    "mg_graph_max_degree_vertex": (C.c_int, [_P, _I32P]),
    "mg_graph_free": (None, [_P]),
    "mg_default_options": (None, [C.POINTER(Options)]),
    "mg_device_count": (C.c_int, [_I32P]),
    "mg_build": (C.c_int, [_P, C.c_int, C.POINTER(Options), C.POINTER(_P)]),
    "mg_release": (C.c_int, [_P]),
    "mg_reset": (C.c_int, [_P, C.c_int]),
    "mg_step": (C.c_int, [_P, _I32P]),
    "mg_bc_backward": (C.c_int, [_P, C.c_int]),
    "mg_run": (C.c_int, [_P, C.c_int, C.POINTER(C.c_float)]),
    "mg_sync": (C.c_int, [_P]),
    "mg_iterations": (C.c_int, [_P, _I32P]),
    "mg_get_result": (C.c_int, [_P, C.c_int, _P, C.c_int64]),
    "mg_batch_build": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(Options), C.POINTER(_P)]),
    "mg_batch_run": (C.c_int, [_P, C.c_int, _I32P, C.POINTER(C.c_float)]),
    "mg_batch_run_slots": (C.c_int, [_P, C.c_int, _I32P, _I32P, C.POINTER(C.c_float)]),
    "mg_batch_get": (C.c_int, [_P, C.c_int, _P, C.c_int64]),
    "mg_batch_release": (C.c_int, [_P]),
    "mg_batch_levels": (C.c_int, [_P, _I32P]),
    "mg_batch_stats": (C.c_int, [_P, C.POINTER(C.c_int64), C.c_int]),
    "mg_step_mode": (C.c_int, [_P, _I32P]),
    "mg_autotune": (C.c_int, [_P, C.c_int, C.c_int, _I32P, C.c_int, C.POINTER(Options), C.POINTER(C.c_float),
                              C.POINTER(C.c_float)]),
    "mg_batch_autotune": (C.c_int, [_P, C.c_int, C.c_int, _I32P, _I32P, C.c_int, C.POINTER(Options),
                                    C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "mg_bfs_noalloc": (C.c_int, [C.c_int, C.c_int64, _P, _P, C.c_int, C.c_int, _P, _P]),
    "mg_set_profiling": (C.c_int, [_P, C.c_int]),
    "mg_set_profile_filter": (C.c_int, [_P, C.c_char_p]),
    "mg_kernel_stats": (C.c_int, [_P, C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]),
}


def lib():
    """Load libmgamd.so (once); raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C {_HERE}` or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc: int) -> int:
    if rc < 0:
        raise MGError(rc, lib().mg_last_error().decode(errors="replace"))
    return rc


def default_options(**kw) -> Options:
    o = Options()
    lib().mg_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def device_count() -> int:
    n = C.c_int32(0)
    rc = lib().mg_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _ptr(a: np.ndarray, ctype=C.c_int32):
    return a.ctypes.data_as(C.POINTER(ctype))


class Graph:
    """Host CSR after reading (the reference's struct csc_package)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    def __del__(self):
        try:
            if self._h and self._h.value and _lib is not None:
                _lib.mg_graph_free(self._h)
                self._h = C.c_void_p(None)
        except Exception:
            pass

    # ---- constructors
    @classmethod
    def load_mtx(cls, path: str, source: int = 0, flags: int = READ_SYM, weight_seed: int = 1) -> "Graph":
        h = C.c_void_p()
        _check(lib().mg_graph_load_mtx(str(path).encode(), source, flags, weight_seed, C.byref(h)))
        return cls(h)

    @classmethod
    def from_coo(cls, nv: int, src, dst, val=None, source: int = 0, flags: int = READ_SYM) -> "Graph":
        s, d = _i32(src), _i32(dst)
        v = _i32(val) if val is not None else None
        h = C.c_void_p()
        _check(lib().mg_graph_from_coo(nv, len(s), _ptr(s), _ptr(d), _ptr(v) if v is not None else None, source, flags,
                                       C.byref(h)))
        return cls(h)

    @classmethod
    def from_csr(cls, row, col, weight=None) -> "Graph":
        """a CSR built by the caller: ids kept, rows sorted by column, repeats dropped"""
        r = np.ascontiguousarray(row, dtype=np.int64)
        c = _i32(col)
        w = _i32(weight) if weight is not None else None
        h = C.c_void_p()
        _check(lib().mg_graph_from_csr(len(r) - 1, r.ctypes.data_as(C.POINTER(C.c_int64)), _ptr(c),
                                       _ptr(w) if w is not None else None, C.byref(h)))
        return cls(h)

    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19, seed: int = 1,
             weight_seed: int = 7, flags: int = READ_SYM) -> "Graph":
        h = C.c_void_p()
        _check(lib().mg_graph_gen_rmat(scale, edge_factor, a, b, c, seed, weight_seed, flags, C.byref(h)))
        return cls(h)

    @classmethod
    def grid(cls, width: int, height: int, weight_seed: int = 7, flags: int = READ_SYM) -> "Graph":
        h = C.c_void_p()
        _check(lib().mg_graph_gen_grid(width, height, weight_seed, flags, C.byref(h)))
        return cls(h)

    @classmethod
    def union(cls, graphs) -> "Graph":
        arr = (C.c_void_p * len(graphs))(*[g._h.value for g in graphs])
        h = C.c_void_p()
        _check(lib().mg_graph_union(len(graphs), arr, C.byref(h)))
        return cls(h)

    # ---- accessors
    def info(self):
        nv, ne, sp = C.c_int32(), C.c_int64(), C.c_int32()
        _check(lib().mg_graph_info(self._h, C.byref(nv), C.byref(ne), C.byref(sp)))
        return nv.value, ne.value, sp.value

    @property
    def symmetric(self) -> bool:
        """every stored edge has its reverse with the same weight"""
        v = C.c_int32()
        _check(lib().mg_graph_symmetric(self._h, C.byref(v)))
        return bool(v.value)

    @property
    def nv(self) -> int:
        return self.info()[0]

    @property
    def ne(self) -> int:
        return self.info()[1]

    @property
    def start_p(self) -> int:
        return self.info()[2]

    def csr(self, copy: bool = True):
        """(row, col, weight) as numpy arrays: copies, or (copy=False) views of the graph's
        own col / weight arrays, valid while the graph lives.  Row offsets are int32, or
        int64 on a graph past 2^31 - 1 edges."""
        nv, ne, _ = self.info()
        r, c, w = _I32P(), _I32P(), _I32P()
        if ne > 2**31 - 1:
            row = np.empty(nv + 1, np.int64)
            _check(lib().mg_graph_csr64(self._h, row.ctypes.data_as(C.POINTER(C.c_int64))))
            _check(lib().mg_graph_csr(self._h, None, C.byref(c), C.byref(w)))
        else:
            _check(lib().mg_graph_csr(self._h, C.byref(r), C.byref(c), C.byref(w)))
            row = np.ctypeslib.as_array(r, shape=(nv + 1,)).copy()
        col = np.ctypeslib.as_array(c, shape=(ne,)) if ne else np.zeros(0, np.int32)
        wt = np.ctypeslib.as_array(w, shape=(ne,)) if ne else np.zeros(0, np.int32)
        if copy:
            col, wt = col.copy(), wt.copy()
        return row, col, wt

    def orig_ids(self) -> np.ndarray:
        nv = self.nv
        o = _I32P()
        _check(lib().mg_graph_orig_ids(self._h, C.byref(o)))
        return np.ctypeslib.as_array(o, shape=(nv,)).copy() if nv else np.zeros(0, np.int32)

    def remap(self, orig_id: int) -> int:
        v = C.c_int32()
        _check(lib().mg_graph_remap(self._h, orig_id, C.byref(v)))
        return v.value

    def max_degree_vertex(self) -> int:
        v = C.c_int32()
        _check(lib().mg_graph_max_degree_vertex(self._h, C.byref(v)))
        return v.value

    def write_mtx(self, path: str) -> None:
        _check(lib().mg_graph_write_mtx(self._h, str(path).encode()))


class Engine:
    """Device layout + run state of one algorithm on one graph (HYB + vector_data)."""

    def __init__(self, graph: Graph, algo: int, options: Options | None = None, **kw):
        self.graph = graph
        self.algo = algo
        self.opt = options if options is not None else default_options(**kw)
        h = C.c_void_p()
        _check(lib().mg_build(graph._h, algo, C.byref(self.opt), C.byref(h)))
        self._h = h
        self.nv = graph.nv

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.mg_release(self._h)
            self._h = C.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self, source: int = 0) -> None:
        _check(lib().mg_reset(self._h, source))

    def step(self) -> tuple[int, int]:
        it = C.c_int32(0)
        more = _check(lib().mg_step(self._h, C.byref(it)))
        return more, it.value

    def step_mode(self) -> int:
        """mode of the level the last step ran: 0 dense (pull), 1 sparse (push), -1 none"""
        m = C.c_int32(-1)
        _check(lib().mg_step_mode(self._h, C.byref(m)))
        return m.value

    def bc_backward(self) -> None:
        _check(lib().mg_bc_backward(self._h, 0))

    def run(self, source: int = 0) -> float:
        ms = C.c_float(0.0)
        _check(lib().mg_run(self._h, source, C.byref(ms)))
        return ms.value

    def iterations(self) -> int:
        n = C.c_int32()
        _check(lib().mg_iterations(self._h, C.byref(n)))
        return n.value

    def result(self, which: int = RES_LEVEL) -> np.ndarray:
        dt = {RES_LEVEL: np.int32, RES_SIGMA: np.float64, RES_DELTA: np.float32, RES_RANK: np.float32}[which]
        out = np.empty(self.nv, dtype=dt)
        _check(lib().mg_get_result(self._h, which, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def set_profiling(self, on: bool, only: str | None = None) -> None:
        """per-kernel HIP-event accounting; `only`: a kernel-name prefix to restrict it to"""
        _check(lib().mg_set_profile_filter(self._h, only.encode() if only else None))
        _check(lib().mg_set_profiling(self._h, 1 if on else 0))

    def kernel_stats(self) -> dict:
        out = {}
        i = 0
        name = C.create_string_buffer(128)
        while True:
            n, ms, by = C.c_int64(), C.c_double(), C.c_double()
            rc = lib().mg_kernel_stats(self._h, i, name, 128, C.byref(n), C.byref(ms), C.byref(by))
            if rc < 0:
                break
            out[name.value.decode()] = {"launches": n.value, "ms": ms.value, "bytes": by.value}
            i += 1
        return out


class Batch:
    """Batched multi-source runs: one bit per source packed per vertex."""

    def __init__(self, graph: Graph, kind: int, max_sources: int = 64, options: Options | None = None, **kw):
        self.graph = graph
        self.kind = kind
        self.nv = graph.nv
        self.opt = options if options is not None else default_options(**kw)
        h = C.c_void_p()
        _check(lib().mg_batch_build(graph._h, kind, max_sources, C.byref(self.opt), C.byref(h)))
        self._h = h
        self.nsrc = 0

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.mg_batch_release(self._h)
            self._h = C.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, sources, slots=None) -> float:
        """one run; source i takes bit i, or bit slots[i] when a slot list is given"""
        s = _i32(sources)
        ms = C.c_float(0.0)
        if slots is None:
            _check(lib().mg_batch_run(self._h, len(s), _ptr(s), C.byref(ms)))
            self.nsrc = len(s)
        else:
            k = _i32(slots)
            if len(k) != len(s):
                raise ValueError("one slot per source")
            _check(lib().mg_batch_run_slots(self._h, len(s), _ptr(s), _ptr(k), C.byref(ms)))
            self.nsrc = int(k.max()) + 1
        return ms.value

    def levels(self) -> np.ndarray:
        """[nsrc][nv] int32: BFS/BC levels, or SSSP distances for an SSSP batch"""
        out = np.empty((self.nsrc, self.nv), dtype=np.int32)
        _check(lib().mg_batch_get(self._h, RES_LEVEL, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def sigma(self) -> np.ndarray:
        out = np.empty((self.nsrc, self.nv), dtype=np.float64)
        _check(lib().mg_batch_get(self._h, RES_SIGMA, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def sweeps(self) -> int:
        n = C.c_int32()
        _check(lib().mg_batch_levels(self._h, C.byref(n)))
        return n.value

    def stats(self) -> dict:
        """how the last BFS / SSSP run spent its levels"""
        out = (C.c_int64 * 5)()
        _check(lib().mg_batch_stats(self._h, out, 5))
        return {"levels": out[0], "dense_levels": out[1], "sparse_levels": out[2], "persistent_levels": out[3],
                "host_round_trips": out[4]}

    def bc(self) -> np.ndarray:
        out = np.empty(self.nv, dtype=np.float32)
        _check(lib().mg_batch_get(self._h, RES_DELTA, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out


def autotune(graph: Graph, algo: int, sources, reps: int = 3, **kw):
    """(best Options, ms with the defaults, ms tuned) for an engine algorithm"""
    src = _i32(sources)
    o = default_options(**kw)
    a, b = C.c_float(), C.c_float()
    _check(lib().mg_autotune(graph._h, algo, len(src), _ptr(src), reps, C.byref(o), C.byref(a), C.byref(b)))
    return o, a.value, b.value


def batch_autotune(graph: Graph, kind: int, sources, slots=None, reps: int = 3, **kw):
    """(best Options, ms with the defaults, ms tuned) for one batched run"""
    src = _i32(sources)
    sl = _i32(slots) if slots is not None else None
    o = default_options(**kw)
    a, b = C.c_float(), C.c_float()
    _check(lib().mg_batch_autotune(graph._h, kind, len(src), _ptr(src), _ptr(sl) if sl is not None else None, reps,
                                   C.byref(o), C.byref(a), C.byref(b)))
    return o, a.value, b.value


def bfs_noalloc(nv: int, ne: int, d_row: int, d_col: int, source: int, d_level: int, mode: int = MODE_HYBRID,
                stream: int = 0) -> None:
    """BFS levels into the device buffer d_level (nv int32) over the device CSR d_row / d_col
    (int32), with no device allocation: the library's fixed scratch array holds the run state.
    Pointers are device addresses as integers; calls on one device must not overlap.  Pull
    levels read rows as in-edges: a directed CSR needs mode=MODE_SPARSE."""
    _check(lib().mg_bfs_noalloc(nv, ne, C.c_void_p(d_row), C.c_void_p(d_col), source, mode, C.c_void_p(d_level),
                                C.c_void_p(stream) if stream else None))
