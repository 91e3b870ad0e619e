"""CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package; the product never does.  liboracle.so is the C restatement of the
reference CPU algorithms in mg_oracle.c (built by oracle/Makefile).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None
I32P = C.POINTER(C.c_int32)


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists() or LIB.stat().st_mtime < (HERE / "mg_oracle.c").stat().st_mtime:
            build()
        L = C.CDLL(str(LIB))
        L.oracle_bfs.argtypes = [C.c_int, C.c_int, I32P, I32P, I32P, I32P, C.c_int, I32P]
        L.oracle_sssp_quadratic.argtypes = [C.c_int, C.c_int, I32P, I32P, I32P, I32P, I32P, I32P, C.c_int,
                                            C.c_int, I32P]
        L.oracle_sssp_quadratic.restype = None
        L.oracle_sssp_heap.argtypes = L.oracle_sssp_quadratic.argtypes
        L.oracle_sssp_heap.restype = None
        L.oracle_cc.argtypes = [C.c_int, I32P, I32P, I32P, I32P, I32P]
        L.oracle_cc_union_find.argtypes = [C.c_int, I32P, I32P, I32P]
        L.oracle_cc_union_find.restype = None
        L.oracle_pr_gauss_seidel.argtypes = [C.c_int, I32P, I32P, I32P, C.c_int, C.POINTER(C.c_float)]
        L.oracle_pr_gas.argtypes = [C.c_int, I32P, I32P, I32P, I32P, C.c_int, C.c_float, C.c_float,
                                    C.POINTER(C.c_float)]
        L.oracle_pr_jacobi_f32.argtypes = [C.c_int, I32P, I32P, I32P, C.c_int, C.c_float, C.c_double,
                                           C.POINTER(C.c_float), C.POINTER(C.c_double)]
        L.oracle_pr_jacobi_f64.argtypes = [C.c_int, I32P, I32P, I32P, C.c_int, C.c_double, C.c_double,
                                           C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.oracle_teps_edges.argtypes = [C.c_int, I32P, I32P, C.c_int]
        L.oracle_teps_edges.restype = C.c_longlong
        _lib = L
    return _lib


def _p(a, t=C.c_int32):
    return None if a is None else np.ascontiguousarray(a).ctypes.data_as(C.POINTER(t))


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _csc(row_offsets, col_idx):
    """Stable CSC (col_offsets, row_idx, csr_edge_id) computed with numpy."""
    n = len(row_offsets) - 1
    rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(row_offsets))
    order = np.argsort(col_idx, kind="stable").astype(np.int32)
    co = np.zeros(n + 1, np.int64)
    np.add.at(co, np.asarray(col_idx, np.int64) + 1, 1)
    return np.cumsum(co).astype(np.int32), rows[order], order


def bfs(row_offsets, col_idx, src, directed=1, csc=None):
    ro, ci = _i32(row_offsets), _i32(col_idx)
    n = len(ro) - 1
    co, ri, _ = csc if csc is not None else (_csc(ro, ci) if directed == 0 else (None, None, None))
    out = np.empty(n, np.int32)
    lib().oracle_bfs(directed, n, _p(ro), _p(ci), _p(_i32(co)), _p(_i32(ri)), src, _p(out))
    return out


def sssp(row_offsets, col_idx, weights, src, directed=1, unreached=2**31 - 1, quadratic=False):
    ro, ci, w = _i32(row_offsets), _i32(col_idx), _i32(weights)
    n = len(ro) - 1
    co = ri = ce = None
    if directed == 0:
        co, ri, ce = _csc(ro, ci)
    out = np.empty(n, np.int32)
    fn = lib().oracle_sssp_quadratic if quadratic else lib().oracle_sssp_heap
    fn(directed, n, _p(ro), _p(ci), _p(w), _p(co), _p(ri), _p(ce), src, unreached, _p(out))
    return out


def cc(row_offsets, col_idx):
    ro, ci = _i32(row_offsets), _i32(col_idx)
    n = len(ro) - 1
    co, ri, _ = _csc(ro, ci)
    out = np.empty(n, np.int32)
    lib().oracle_cc(n, _p(ro), _p(ci), _p(co), _p(ri), _p(out))
    return out


def cc_union_find(row_offsets, col_idx):
    """The labels of cc() (min vertex id per weakly connected component) by union-find;
    fast enough for the scale-22 graphs."""
    ro, ci = _i32(row_offsets), _i32(col_idx)
    out = np.empty(len(ro) - 1, np.int32)
    lib().oracle_cc_union_find(len(ro) - 1, _p(ro), _p(ci), _p(out))
    return out


def pr_gauss_seidel(row_offsets, col_idx, max_sweeps=10000):
    ro, ci = _i32(row_offsets), _i32(col_idx)
    n = len(ro) - 1
    co, ri, _ = _csc(ro, ci)
    out = np.empty(n, np.float32)
    it = lib().oracle_pr_gauss_seidel(n, _p(ro), _p(co), _p(ri), max_sweeps, _p(out, C.c_float))
    return out, it


def pr_gas(row_offsets, col_idx, iter_num=2**31 - 1, teleport=0.15, eps=0.01):
    ro, ci = _i32(row_offsets), _i32(col_idx)
    n = len(ro) - 1
    co, ri, _ = _csc(ro, ci)
    out = np.empty(n, np.float32)
    it = lib().oracle_pr_gas(n, _p(ro), _p(ci), _p(co), _p(ri), iter_num, teleport, eps, _p(out, C.c_float))
    return out, it


def pr_jacobi_f32(row_offsets, col_idx, iters=50, teleport=0.15, tol=0.0):
    ro, ci = _i32(row_offsets), _i32(col_idx)
    n = len(ro) - 1
    co, ri, _ = _csc(ro, ci)
    out = np.empty(n, np.float32)
    l1 = C.c_double(0)
    it = lib().oracle_pr_jacobi_f32(n, _p(ro), _p(co), _p(ri), iters, teleport, tol, _p(out, C.c_float),
                                    C.byref(l1))
    return out, it, l1.value


def pr_jacobi_f64(row_offsets, col_idx, iters=50, teleport=0.15, tol=0.0):
    ro, ci = _i32(row_offsets), _i32(col_idx)
    n = len(ro) - 1
    co, ri, _ = _csc(ro, ci)
    out = np.empty(n, np.float64)
    l1 = C.c_double(0)
    it = lib().oracle_pr_jacobi_f64(n, _p(ro), _p(co), _p(ri), iters, teleport, tol, _p(out, C.c_double),
                                    C.byref(l1))
    return out, it, l1.value


def teps_edges(row_offsets, labels, unreached=-1):
    ro, lb = _i32(row_offsets), _i32(labels)
    return int(lib().oracle_teps_edges(len(ro) - 1, _p(ro), _p(lb), unreached))


def ppr(row_offsets, col_idx, src, iters=20, teleport=0.15):
    """Personalized PageRank in float64: x_0 = e_src, x_{i+1} = t e_src + (1 - t) A^T D^-1 x_i over the
    CSR edges v -> u (rank of out-degree-0 vertices is dropped, as in PageRank.h:233-283)."""
    ro, ci = np.asarray(row_offsets, np.int64), np.asarray(col_idx, np.int64)
    n = len(ro) - 1
    deg = np.diff(ro)
    rows = np.repeat(np.arange(n), deg)
    x = np.zeros(n)
    x[src] = 1.0
    inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    for _ in range(iters):
        y = np.bincount(ci, weights=(x * inv)[rows], minlength=n) * (1.0 - teleport)
        y[src] += teleport
        x = y
    return x
