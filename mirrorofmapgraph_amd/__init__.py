"""Python binding of libmapgraph_hip.so (ctypes over the C-ABI in include/mapgraph_c.h).

The engine itself is HIP/C++; this module only marshals numpy arrays across the
C-ABI.  It never falls back to a CPU implementation: loading fails loudly when the
shared library has not been built, and every search runs on the GPU.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libmapgraph_hip.so"

MG_OK = 0
MG_ERR_INVALID_ARG = 1
MG_ERR_OOM = 2
MG_ERR_DEVICE = 3
MG_ERR_OVERFLOW = 4
MG_ERR_IO = 5
MG_ERR_FORMAT = 6
MG_ERR_UNSUPPORTED = 7
MG_ERR_NO_DEVICE = 8
MG_ERR_TIMEOUT = 9
MG_Q_BFS, MG_Q_SSSP, MG_Q_PPR = 0, 1, 2
MG_BATCH_MAX = 256
MG_PPR_MAX = 64

MG_GRAPH_WEIGHTED = 0x1
MG_GRAPH_BUILD_CSC = 0x2
MG_GRAPH_SYMMETRIC = 0x4
MG_GRAPH_RELABEL = 0x8

MG_ENGINE_AUTO = 0
MG_ENGINE_GAS = 1
MG_PR_GAS = 0
MG_PR_ALL = 1

MG_CSR_SYMMETRIZE = 0x1
MG_CSR_DROP_LOOPS = 0x2
MG_CSR_DEDUPE = 0x4
MG_CSR_SORT_COLS = 0x8

INT32_MAX = 2**31 - 1

# Symbols every build must export (checked by the CPU test suite).
EXPORTED = [
    "mg_last_error", "mg_version", "mg_device_init", "mg_device_count", "mg_device_synchronize",
    "mg_graph_create", "mg_graph_destroy", "mg_graph_info", "mg_graph_device_bytes", "mg_graph_get_csc",
    "mg_options_default", "mg_bfs", "mg_bfs_pred", "mg_sssp", "mg_cc", "mg_pagerank", "mg_bfs_batch", "mg_sssp_batch", "mg_ppr_batch",
    "mg_rmat_num_edges", "mg_gen_rmat_edges", "mg_grid_num_edges", "mg_gen_grid_edges",
    "mg_csr_from_edges", "mg_edge_weights_hash", "mg_csc_from_csr", "mg_pick_sources",
    "mg_graph_union", "mg_bfs_batch_sets", "mg_sssp_batch_sets",
    "mg_mtx_read_header", "mg_mtx_read", "mg_autotune", "mg_csr_relabel_degree",
    "mg_server_config_default", "mg_server_create", "mg_server_add_graph", "mg_server_submit", "mg_server_wait",
    "mg_server_now_ms", "mg_server_get_stats", "mg_server_destroy", "mg_host_alloc", "mg_host_free",
]


class MgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[mg status {code}] {msg}")
        self.code = code


class Options(C.Structure):
    _fields_ = [
        ("iter_num", C.c_int32), ("threshold", C.c_int32), ("directed", C.c_int32),
        ("engine", C.c_int32), ("verbose", C.c_int32), ("out_on_device", C.c_int32),
        ("pr_mode", C.c_int32), ("pr_teleport", C.c_float), ("pr_tol", C.c_float),
        ("pr_active_eps", C.c_float), ("max_queue_sizing", C.c_double),
        ("alpha", C.c_int32), ("beta", C.c_int32), ("batch_alpha", C.c_int32), ("batch_beta", C.c_int32),
        ("sssp_delta", C.c_int32), ("ppr_iters", C.c_int32), ("sssp_bucket_edges", C.c_int32),
        ("level_grid", C.c_int32), ("layout", C.c_int32), ("reserved", C.c_int32 * 1),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("edges_traversed", C.c_int64), ("vertices_reached", C.c_int64),
        ("edges_examined", C.c_int64), ("iterations", C.c_int32),
        ("overflow_regrows", C.c_int32), ("device_ms", C.c_double), ("host_ms", C.c_double),
        ("bottom_up_levels", C.c_int32), ("reserved", C.c_int32 * 7),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_lib = None
class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("max_slots", C.c_int32), ("ppr_iters", C.c_int32),
                ("ppr_teleport", C.c_float), ("steps_per_round", C.c_int32), ("sssp_delta", C.c_int32),
                ("thin_levels", C.c_int32), ("thin_lanes", C.c_int32), ("thin_streams", C.c_int32),
                ("thin_slice_us", C.c_int32), ("ppr_gate_us", C.c_int32),
                ("ppr_batch_min", C.c_int32), ("plain_kinds", C.c_int32)]


MG_TUNE_BFS, MG_TUNE_SSSP, MG_TUNE_BFS_BATCH, MG_TUNE_SSSP_BATCH = 0, 1, 2, 3
TUNE_KINDS = {"bfs": MG_TUNE_BFS, "sssp": MG_TUNE_SSSP, "bfs_batch": MG_TUNE_BFS_BATCH,
              "sssp_batch": MG_TUNE_SSSP_BATCH}
# the mg_options fields mg_autotune chooses
TUNABLE = ("alpha", "beta", "level_grid", "sssp_delta", "sssp_bucket_edges", "batch_alpha", "batch_beta")


class TuneReport(C.Structure):
    _fields_ = [("start_ms", C.c_double), ("best_ms", C.c_double), ("candidates", C.c_int32),
                ("reserved", C.c_int32 * 7)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class QueryInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("kind", C.c_int32), ("graph", C.c_int32), ("src", C.c_int32),
                ("edges_traversed", C.c_int64), ("vertices_reached", C.c_int64), ("t_submit_ms", C.c_double),
                ("t_admit_ms", C.c_double), ("t_done_ms", C.c_double), ("reserved", C.c_int32 * 4)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("rounds", C.c_int64),
                ("edges_traversed", C.c_int64), ("device_bytes", C.c_int64), ("latency_ms_sum", C.c_double),
                ("thin_engines", C.c_int32), ("pad", C.c_int32), ("dispatch_busy_ms", C.c_double),
                ("dispatch_passes", C.c_int64), ("reserved", C.c_int32 * 2)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


I32P = C.POINTER(C.c_int32)
F32P = C.POINTER(C.c_float)


def _p(a: np.ndarray | None, ctype=C.c_int32):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def lib() -> C.CDLL:
    """Loads the built library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m mirrorofmapgraph_amd.build`")
    L = C.CDLL(str(LIB_PATH))
    L.mg_last_error.restype = C.c_size_t
    L.mg_last_error.argtypes = [C.c_char_p, C.c_size_t]
    L.mg_version.restype = C.c_char_p
    L.mg_rmat_num_edges.restype = C.c_int64
    L.mg_rmat_num_edges.argtypes = [C.c_int32, C.c_int32]
    L.mg_grid_num_edges.restype = C.c_int64
    L.mg_grid_num_edges.argtypes = [C.c_int32, C.c_int32]
    L.mg_gen_rmat_edges.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                    I32P, I32P]
    L.mg_gen_grid_edges.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_uint64, I32P, I32P,
                                    C.POINTER(C.c_int64)]
    L.mg_csr_from_edges.argtypes = [C.c_int32, C.c_int64, I32P, I32P, I32P, C.c_uint32, I32P, I32P, I32P,
                                    C.c_int64, C.POINTER(C.c_int64)]
    L.mg_edge_weights_hash.argtypes = [C.c_int32, I32P, I32P, C.c_int64, C.c_uint64, C.c_int32, I32P]
    L.mg_csc_from_csr.argtypes = [C.c_int32, C.c_int64, I32P, I32P, I32P, I32P, I32P]
    L.mg_pick_sources.argtypes = [C.c_int32, I32P, C.c_int32, C.c_uint64, I32P]
    if getattr(L, "mg_csr_relabel_degree", None) is not None:
        L.mg_csr_relabel_degree.argtypes = [C.c_int32, C.c_int64, I32P, I32P, I32P, I32P, I32P, I32P, I32P]
    L.mg_mtx_read_header.argtypes = [C.c_char_p, I32P, C.POINTER(C.c_int64), I32P]
    L.mg_mtx_read.argtypes = [C.c_char_p, I32P, I32P, I32P, C.c_int64, C.POINTER(C.c_int64)]
    L.mg_graph_create.argtypes = [I32P, I32P, I32P, C.c_int32, C.c_int64, C.c_uint32, C.POINTER(C.c_void_p)]
    L.mg_graph_destroy.argtypes = [C.c_void_p]
    L.mg_graph_info.argtypes = [C.c_void_p, I32P, C.POINTER(C.c_int64), C.POINTER(C.c_uint32)]
    L.mg_graph_device_bytes.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    L.mg_graph_get_csc.argtypes = [C.c_void_p, I32P, I32P, I32P]
    L.mg_options_default.argtypes = [C.POINTER(Options)]
    L.mg_options_default.restype = None
    L.mg_device_init.argtypes = [C.c_int]
    search = {
        "mg_bfs": [C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(Options), C.POINTER(Stats)],
        "mg_bfs_pred": [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.POINTER(Options), C.POINTER(Stats)],
        "mg_sssp": [C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(Options), C.POINTER(Stats)],
        "mg_cc": [C.c_void_p, C.c_void_p, C.POINTER(Options), C.POINTER(Stats)],
        "mg_pagerank": [C.c_void_p, C.c_void_p, C.POINTER(Options), C.POINTER(Stats)],
        "mg_bfs_batch": [C.c_void_p, I32P, C.c_int32, C.c_void_p, C.POINTER(Options), C.POINTER(Stats)],
        "mg_bfs_batch_sets": [C.c_void_p, I32P, I32P, C.c_int32, C.c_int32, C.c_void_p, C.POINTER(Options),
                              C.POINTER(Stats)],
        "mg_sssp_batch_sets": [C.c_void_p, I32P, I32P, C.c_int32, C.c_int32, C.c_void_p, C.POINTER(Options),
                               C.POINTER(Stats)],
        "mg_graph_union": [C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), I32P],
        "mg_sssp_batch": [C.c_void_p, I32P, C.c_int32, C.c_void_p, C.POINTER(Options), C.POINTER(Stats)],
        "mg_ppr_batch": [C.c_void_p, I32P, C.c_int32, C.c_void_p, C.POINTER(Options), C.POINTER(Stats)],
    }
    search.update({
        "mg_autotune": [C.c_void_p, C.c_int32, I32P, C.c_int32, C.c_int32, C.POINTER(Options), C.POINTER(Options),
                        C.POINTER(TuneReport)],
        "mg_server_config_default": [C.POINTER(ServerConfig)],
        "mg_server_create": [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)],
        "mg_server_add_graph": [C.c_void_p, C.c_void_p, I32P],
        "mg_server_submit": [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_int32,
                             C.POINTER(C.c_int64)],
        "mg_server_wait": [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(QueryInfo)],
        "mg_server_now_ms": [C.c_void_p, C.POINTER(C.c_double)],
        "mg_server_get_stats": [C.c_void_p, C.POINTER(ServerStats)],
        "mg_server_destroy": [C.c_void_p],
        "mg_host_alloc": [C.c_size_t],
        "mg_host_free": [C.c_void_p],
    })
    for name, argtypes in search.items():
        fn = getattr(L, name, None)  # a missing export is reported by the export test
        if fn is not None:
            fn.argtypes = argtypes
    for name in ("mg_server_config_default", "mg_host_free"):
        if getattr(L, name, None) is not None:
            getattr(L, name).restype = None
    if getattr(L, "mg_host_alloc", None) is not None:
        L.mg_host_alloc.restype = C.c_void_p
    _lib = L
    return L


def last_error() -> str:
    L = lib()
    buf = C.create_string_buffer(2048)
    L.mg_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int) -> None:
    if rc != MG_OK:
        raise MgError(rc, last_error())


def options(**kw) -> Options:
    o = Options()
    lib().mg_options_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(f"unknown option {k}")
        setattr(o, k, v)
    return o


def device_count() -> int:
    return int(lib().mg_device_count())


# ------------------------------------------------------------------ host graph tools
class HostCSR:
    """A host CSR graph (int32 row_offsets / col_idx, optional int32 values)."""

    def __init__(self, row_offsets: np.ndarray, col_idx: np.ndarray, values: np.ndarray | None = None):
        self.row_offsets = np.ascontiguousarray(row_offsets, dtype=np.int32)
        self.col_idx = np.ascontiguousarray(col_idx, dtype=np.int32)
        self.values = None if values is None else np.ascontiguousarray(values, dtype=np.int32)

    @property
    def n(self) -> int:
        return len(self.row_offsets) - 1

    @property
    def m(self) -> int:
        return len(self.col_idx)

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_offsets)

    def csc(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(col_offsets, row_idx, csr_edge_id) by a stable counting sort on the column."""
        co = np.empty(self.n + 1, np.int32)
        ri = np.empty(self.m, np.int32)
        ce = np.empty(self.m, np.int32)
        check(lib().mg_csc_from_csr(self.n, self.m, _p(self.row_offsets), _p(self.col_idx), _p(co), _p(ri),
                                    _p(ce)))
        return co, ri, ce


def rmat_edges(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
               seed: int = 42) -> tuple[np.ndarray, np.ndarray]:
    m = int(lib().mg_rmat_num_edges(scale, edge_factor))
    src = np.empty(m, np.int32)
    dst = np.empty(m, np.int32)
    check(lib().mg_gen_rmat_edges(scale, edge_factor, a, b, c, seed, _p(src), _p(dst)))
    return src, dst


def grid_edges(width: int, height: int, delete_frac: float = 0.05, seed: int = 42) -> tuple[np.ndarray, np.ndarray]:
    E = int(lib().mg_grid_num_edges(width, height))
    src = np.empty(E, np.int32)
    dst = np.empty(E, np.int32)
    m = C.c_int64(0)
    check(lib().mg_gen_grid_edges(width, height, delete_frac, seed, _p(src), _p(dst), C.byref(m)))
    return src[: m.value].copy(), dst[: m.value].copy()


def csr_from_edges(n: int, src: np.ndarray, dst: np.ndarray, flags: int, vals: np.ndarray | None = None) -> HostCSR:
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    cap = 2 * len(src) if flags & MG_CSR_SYMMETRIZE else len(src)
    ro = np.empty(n + 1, np.int32)
    col = np.empty(max(cap, 1), np.int32)
    vo = None
    if vals is not None:
        vals = np.ascontiguousarray(vals, np.int32)
        vo = np.empty(max(cap, 1), np.int32)
    m = C.c_int64(0)
    check(lib().mg_csr_from_edges(n, len(src), _p(src), _p(dst), _p(vals), flags, _p(ro), _p(col), _p(vo),
                                  cap, C.byref(m)))
    mm = m.value
    return HostCSR(ro, col[:mm].copy(), None if vo is None else vo[:mm].copy())


def hash_weights(g: HostCSR, seed: int, wmax: int) -> np.ndarray:
    w = np.empty(g.m, np.int32)
    check(lib().mg_edge_weights_hash(g.n, _p(g.row_offsets), _p(g.col_idx), g.m, seed, wmax, _p(w)))
    return w


def relabel_degree(g: HostCSR) -> tuple[HostCSR, np.ndarray]:
    """The graph renamed in out-degree order (highest first, ties by id) with every row's
    targets ascending, and new_of_old (new id of each vertex)."""
    perm = np.empty(g.n, np.int32)
    ro = np.empty(g.n + 1, np.int32)
    col = np.empty(max(g.m, 1), np.int32)
    vals = None if g.values is None else np.empty(max(g.m, 1), np.int32)
    check(lib().mg_csr_relabel_degree(g.n, g.m, _p(g.row_offsets), _p(g.col_idx), _p(g.values), _p(perm), _p(ro),
                                      _p(col), _p(vals)))
    return HostCSR(ro, col[: g.m], None if vals is None else vals[: g.m]), perm


def pick_sources(g: HostCSR, k: int, seed: int) -> np.ndarray:
    out = np.empty(k, np.int32)
    check(lib().mg_pick_sources(g.n, _p(g.row_offsets), k, seed, _p(out)))
    return out


def read_mtx(path: str | os.PathLike) -> tuple[int, np.ndarray, np.ndarray, np.ndarray, bool]:
    """(n, src, dst, val, symmetric_flag) of a Matrix Market file (0-based ids)."""
    n = C.c_int32(0)
    nnz = C.c_int64(0)
    sym = C.c_int32(0)
    bp = str(path).encode()
    check(lib().mg_mtx_read_header(bp, C.byref(n), C.byref(nnz), C.byref(sym)))
    src = np.empty(max(nnz.value, 1), np.int32)
    dst = np.empty(max(nnz.value, 1), np.int32)
    val = np.empty(max(nnz.value, 1), np.int32)
    got = C.c_int64(0)
    check(lib().mg_mtx_read(bp, _p(src), _p(dst), _p(val), nnz.value, C.byref(got)))
    k = got.value
    return n.value, src[:k].copy(), dst[:k].copy(), val[:k].copy(), bool(sym.value)


def rmat_graph(scale: int, edge_factor: int = 16, seed: int = 42, symmetric: bool = True,
               weight_max: int | None = None, weight_seed: int | None = None) -> HostCSR:
    """Seeded R-MAT (0.57, 0.19, 0.19, 0.05): self-loops dropped, deduped, sorted
    rows; symmetrized unless symmetric=False; optional hashed weights in [1, weight_max]."""
    src, dst = rmat_edges(scale, edge_factor, seed=seed)
    flags = MG_CSR_DROP_LOOPS | MG_CSR_DEDUPE | MG_CSR_SORT_COLS
    if symmetric:
        flags |= MG_CSR_SYMMETRIZE
    g = csr_from_edges(1 << scale, src, dst, flags)
    del src, dst
    if weight_max:
        g.values = hash_weights(g, seed + 1 if weight_seed is None else weight_seed, weight_max)
    return g


def grid_graph(width: int, height: int | None = None, delete_frac: float = 0.05, seed: int = 42,
               weight_max: int | None = None, weight_seed: int | None = None) -> HostCSR:
    """Symmetric 4-neighbour grid with exactly floor(delete_frac * E) edges removed."""
    height = width if height is None else height
    src, dst = grid_edges(width, height, delete_frac, seed)
    g = csr_from_edges(width * height, src, dst, MG_CSR_SYMMETRIZE | MG_CSR_SORT_COLS)
    if weight_max:
        g.values = hash_weights(g, seed + 1 if weight_seed is None else weight_seed, weight_max)
    return g


# ------------------------------------------------------------------ device graphs
class Graph:
    """A graph resident in device memory (mg_graph)."""

    def __init__(self, host: HostCSR, weighted: bool | None = None, build_csc: bool = False,
                 symmetric: bool = False, relabel: bool = False):
        L = lib()
        self.host = host
        weighted = host.values is not None if weighted is None else weighted
        flags = (MG_GRAPH_WEIGHTED if weighted else 0) | (MG_GRAPH_BUILD_CSC if build_csc else 0) | \
                (MG_GRAPH_SYMMETRIC if symmetric else 0) | (MG_GRAPH_RELABEL if relabel else 0)
        h = C.c_void_p()
        check(L.mg_graph_create(_p(host.row_offsets), _p(host.col_idx), _p(host.values) if weighted else None,
                                host.n, host.m, flags, C.byref(h)))
        self._h = h
        self.n = host.n
        self.m = host.m
        self.flags = flags

    @classmethod
    def union(cls, parts: list["Graph"]) -> "Graph":
        """The disjoint union of resident graphs (mg_graph_union); .offsets[i] is part i's first vertex."""
        L = lib()
        hs = (C.c_void_p * len(parts))(*[p.handle for p in parts])
        offs = np.empty(len(parts) + 1, np.int32)
        h = C.c_void_p()
        check(L.mg_graph_union(len(parts), hs, C.byref(h), _p(offs)))
        g = cls.__new__(cls)
        g.host = None
        g._h = h
        n, m, flags = C.c_int32(), C.c_int64(), C.c_uint32()
        check(L.mg_graph_info(h, C.byref(n), C.byref(m), C.byref(flags)))
        g.n, g.m, g.flags = n.value, m.value, flags.value
        g.offsets = offs
        g.part_n = [p.n for p in parts]
        return g

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().mg_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self) -> C.c_void_p:
        if not self._h:
            raise RuntimeError("graph is closed")
        return self._h

    def device_bytes(self) -> int:
        b = C.c_int64(0)
        check(lib().mg_graph_device_bytes(self.handle, C.byref(b)))
        return b.value

    def csc(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        co = np.empty(self.n + 1, np.int32)
        ri = np.empty(max(self.m, 1), np.int32)
        ce = np.empty(max(self.m, 1), np.int32)
        check(lib().mg_graph_get_csc(self.handle, _p(co), _p(ri), _p(ce)))
        return co, ri[: self.m], ce[: self.m]

    def _run(self, fn, *args, opt: Options | None, out: np.ndarray | None, stats: bool):
        """out=None leaves the result on the device (no copy back)."""
        st = Stats()
        o = opt if opt is not None else options()
        ptr = None if out is None else out.ctypes.data_as(C.c_void_p)
        check(fn(self.handle, *args, ptr, C.byref(o), C.byref(st)))
        return (out, st) if stats else out

    def _out(self, copy: bool, dtype=np.int32):
        return np.empty(self.n, dtype) if copy else None

    def bfs(self, src: int, opt: Options | None = None, stats: bool = False, copy: bool = True):
        return self._run(lib().mg_bfs, src, opt=opt, out=self._out(copy), stats=stats)

    def bfs_pred(self, src: int, opt: Options | None = None, stats: bool = False):
        """(depths, predecessors) of a BFS from src (predextract.cuh semantics)."""
        d = np.empty(self.n, np.int32)
        p = np.empty(self.n, np.int32)
        st = Stats()
        o = opt if opt is not None else options()
        check(lib().mg_bfs_pred(self.handle, src, d.ctypes.data_as(C.c_void_p), p.ctypes.data_as(C.c_void_p),
                                C.byref(o), C.byref(st)))
        return (d, p, st) if stats else (d, p)

    def sssp(self, src: int, opt: Options | None = None, stats: bool = False, copy: bool = True):
        return self._run(lib().mg_sssp, src, opt=opt, out=self._out(copy), stats=stats)

    def cc(self, opt: Options | None = None, stats: bool = False, copy: bool = True):
        return self._run(lib().mg_cc, opt=opt, out=self._out(copy), stats=stats)

    def pagerank(self, opt: Options | None = None, stats: bool = False, copy: bool = True):
        return self._run(lib().mg_pagerank, opt=opt, out=self._out(copy, np.float32), stats=stats)

    def bfs_batch(self, srcs, opt: Options | None = None, stats: bool = False, copy: bool = True):
        """Depths from every source in srcs (at most MG_BATCH_MAX) as a len(srcs) x n array."""
        s = np.ascontiguousarray(srcs, dtype=np.int32)
        out = np.empty((len(s), self.n), np.int32) if copy else None
        return self._run(lib().mg_bfs_batch, _p(s), len(s), opt=opt, out=out, stats=stats)

    def sssp_batch(self, srcs, opt: Options | None = None, stats: bool = False, copy: bool = True):
        """Distances from every source in srcs (at most MG_BATCH_MAX) as a len(srcs) x n array."""
        s = np.ascontiguousarray(srcs, dtype=np.int32)
        out = np.empty((len(s), self.n), np.int32) if copy else None
        return self._run(lib().mg_sssp_batch, _p(s), len(s), opt=opt, out=out, stats=stats)

    def bfs_batch_sets(self, rows_of, srcs, k: int, opt: Options | None = None, stats: bool = False,
                       copy: bool = True):
        """Depths from source sets: row r (of k) from every srcs[i] with rows_of[i] == r (mg_bfs_batch_sets)."""
        r = np.ascontiguousarray(rows_of, dtype=np.int32)
        s = np.ascontiguousarray(srcs, dtype=np.int32)
        out = np.empty((k, self.n), np.int32) if copy else None
        return self._run(lib().mg_bfs_batch_sets, _p(r), _p(s), len(s), k, opt=opt, out=out, stats=stats)

    def sssp_batch_sets(self, rows_of, srcs, k: int, opt: Options | None = None, stats: bool = False,
                        copy: bool = True):
        """Distances from source sets (mg_sssp_batch_sets); see bfs_batch_sets."""
        r = np.ascontiguousarray(rows_of, dtype=np.int32)
        s = np.ascontiguousarray(srcs, dtype=np.int32)
        out = np.empty((k, self.n), np.int32) if copy else None
        return self._run(lib().mg_sssp_batch_sets, _p(r), _p(s), len(s), k, opt=opt, out=out, stats=stats)

    def parts_batch(self, kind: str, srcs_per_part, opt: Options | None = None, stats: bool = False,
                    copy: bool = True):
        """On a union graph: batched BFS ('bfs') or SSSP ('sssp') from srcs_per_part[i] (sources of part i, local
        ids) in one search; row j of every part carries that part's j-th source.  Returns per-part row arrays
        (len(srcs_per_part[i]) x n_i), or None with copy=False, and the stats when asked."""
        rows, src = [], []
        for i, ss in enumerate(srcs_per_part):
            for j, v in enumerate(ss):
                rows.append(j)
                src.append(int(self.offsets[i]) + int(v))
        k = max(len(ss) for ss in srcs_per_part)
        fn = self.bfs_batch_sets if kind == "bfs" else self.sssp_batch_sets
        res = fn(rows, src, k, opt, stats=True, copy=copy)
        out, st = res
        parts = None
        if copy:
            parts = [out[: len(ss), int(self.offsets[i]): int(self.offsets[i]) + self.part_n[i]]
                     for i, ss in enumerate(srcs_per_part)]
        return (parts, st) if stats else parts

    def autotune(self, what: str, srcs, reps: int = 3, start: Options | None = None) -> tuple[Options, dict]:
        """mg_autotune for `what` in TUNE_KINDS on sample sources srcs: the chosen options and the report
        (start_ms / best_ms: device ms of the starting and the chosen setting, candidates timed)."""
        s = np.ascontiguousarray(srcs, dtype=np.int32)
        best, rep = Options(), TuneReport()
        st = start if start is not None else options()
        check(lib().mg_autotune(self.handle, TUNE_KINDS[what], _p(s), len(s), reps, C.byref(st), C.byref(best),
                                C.byref(rep)))
        return best, rep.as_dict()

    def ppr_batch(self, srcs, opt: Options | None = None, stats: bool = False, copy: bool = True):
        """Personalized PageRank from every source in srcs (at most MG_PPR_MAX) as a float32 len(srcs) x n array."""
        s = np.ascontiguousarray(srcs, dtype=np.int32)
        out = np.empty((len(s), self.n), np.float32) if copy else None
        return self._run(lib().mg_ppr_batch, _p(s), len(s), opt=opt, out=out, stats=stats)


class PinnedArray:
    """A numpy array in pinned host memory (mg_host_alloc), for asynchronous result copies."""

    def __init__(self, n: int, dtype=np.int32):
        self.dtype = np.dtype(dtype)
        self.nbytes = int(n) * self.dtype.itemsize
        self.ptr = lib().mg_host_alloc(max(self.nbytes, 1))
        if not self.ptr:
            raise MemoryError("mg_host_alloc failed")
        buf = (C.c_char * max(self.nbytes, 1)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=int(n))

    def close(self) -> None:
        if self.ptr:
            self.array = None
            lib().mg_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop over resident graphs with continuous batching (mg_server)."""

    def __init__(self, **cfg):
        L = lib()
        c = ServerConfig()
        L.mg_server_config_default(C.byref(c))
        for k, v in cfg.items():
            if not hasattr(c, k):
                raise KeyError(f"unknown server option {k}")
            setattr(c, k, v)
        h = C.c_void_p()
        check(L.mg_server_create(C.byref(c), C.byref(h)))
        self._h = h
        self.config = c
        self.graphs = []

    def add_graph(self, g: Graph) -> int:
        gid = C.c_int32()
        check(lib().mg_server_add_graph(self._h, g.handle, C.byref(gid)))
        self.graphs.append(g)
        return gid.value

    def submit(self, graph_id: int, kind: int, src: int, out: "PinnedArray | None" = None) -> int:
        qid = C.c_int64()
        ptr = C.c_void_p(out.ptr) if out is not None else None
        check(lib().mg_server_submit(self._h, graph_id, kind, src, ptr, 0, C.byref(qid)))
        return qid.value

    def wait(self, qid: int, timeout_ms: int = -1) -> dict:
        info = QueryInfo()
        check(lib().mg_server_wait(self._h, qid, timeout_ms, C.byref(info)))
        return info.as_dict()

    def now_ms(self) -> float:
        t = C.c_double()
        check(lib().mg_server_now_ms(self._h, C.byref(t)))
        return t.value

    def stats(self) -> dict:
        st = ServerStats()
        check(lib().mg_server_get_stats(self._h, C.byref(st)))
        return st.as_dict()

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().mg_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
