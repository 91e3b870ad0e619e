"""numpy restatement of the seeded graph generators (test infrastructure only).

The engine's generators are C++ (mirrorofmapgraph_amd/csrc/graph_host.cpp).  This
file restates their definition independently in numpy so tests can check the C++
output bit for bit at small scales:

* mix64: the splitmix64 finalizer (Steele, Lea, Flood 2014).
* R-MAT edge k, level l: r = mix64(mix64(seed) + (64k + l) * 0x9e3779b97f4a7c15) >> 32;
  quadrant (0,0) if r < Ta, (0,1) if r < Tab, (1,0) if r < Tabc, else (1,1), with
  T* = floor(p * 2^32 + 0.5); bits are appended most significant first.
* grid: horizontal edges (r,c)-(r,c+1) numbered r*(W-1)+c, then vertical edges
  (r,c)-(r+1,c) numbered H*(W-1) + r*W + c; the floor(frac*E) edges first in
  (mix64(mix64(seed) + e*golden), e) order are deleted.
* weights: 1 + mix64(mix64(seed) + ((min(u,v) << 32) | max(u,v)) * golden) mod wmax.
* sources: j-th candidate mix64(mix64(seed) + j*golden) mod n, kept if its out-degree > 0.
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def _thr(p: float) -> int:
    return int(p * 4294967296.0 + 0.5)


def rmat_edges(scale: int, edge_factor: int, a: float, b: float, c: float, seed: int):
    m = edge_factor << scale
    ta, tab, tabc = _thr(a), _thr(a + b), _thr(a + b + c)
    ms = mix64(np.uint64(seed))
    k = np.arange(m, dtype=np.uint64)
    u = np.zeros(m, np.uint64)
    v = np.zeros(m, np.uint64)
    with np.errstate(over="ignore"):
        for lvl in range(scale):
            r = mix64(ms + (k * np.uint64(64) + np.uint64(lvl)) * GOLDEN) >> np.uint64(32)
            i = (r >= np.uint64(tab)).astype(np.uint64)
            j = (((r >= np.uint64(ta)) & (r < np.uint64(tab))) | (r >= np.uint64(tabc))).astype(np.uint64)
            u = (u << np.uint64(1)) | i
            v = (v << np.uint64(1)) | j
    return u.astype(np.int32), v.astype(np.int32)


def grid_edges(width: int, height: int, delete_frac: float, seed: int):
    H = height * (width - 1)
    E = H + (height - 1) * width
    D = int(np.floor(delete_frac * E + 1e-9))
    e = np.arange(E, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = mix64(mix64(np.uint64(seed)) + e * GOLDEN)
    order = np.lexsort((np.arange(E), key))
    keep = np.ones(E, bool)
    keep[order[:D]] = False
    ids = np.arange(E)[keep]
    hor = ids < H
    r, c = ids[hor] // (width - 1), ids[hor] % (width - 1)
    u_h = r * width + c
    f = ids[~hor] - H
    u = np.concatenate([u_h, f])
    v = np.concatenate([u_h + 1, f + width])
    # restore canonical edge order
    order2 = np.argsort(np.concatenate([ids[hor], ids[~hor]]), kind="stable")
    return u[order2].astype(np.int32), v[order2].astype(np.int32)


def csr(n: int, src, dst, symmetrize: bool, drop_loops: bool, dedupe: bool):
    """Row-sorted, column-sorted CSR of an edge list."""
    s = np.asarray(src, np.int64)
    d = np.asarray(dst, np.int64)
    if symmetrize:
        s, d = np.concatenate([s, d]), np.concatenate([d, s])
    if drop_loops:
        keep = s != d
        s, d = s[keep], d[keep]
    key = s * n + d
    key = np.unique(key) if dedupe else np.sort(key, kind="stable")
    s, d = key // n, key % n
    ro = np.zeros(n + 1, np.int64)
    np.add.at(ro, s + 1, 1)
    return np.cumsum(ro).astype(np.int32), d.astype(np.int32)


def weights(row_offsets, col_idx, seed: int, wmax: int):
    n = len(row_offsets) - 1
    u = np.repeat(np.arange(n, dtype=np.uint64), np.diff(row_offsets))
    v = np.asarray(col_idx, np.uint64)
    lo, hi = np.minimum(u, v), np.maximum(u, v)
    with np.errstate(over="ignore"):
        h = mix64(mix64(np.uint64(seed)) + ((lo << np.uint64(32)) | hi) * GOLDEN)
    return (np.uint64(1) + h % np.uint64(wmax)).astype(np.int32)


def sources(row_offsets, k: int, seed: int):
    n = len(row_offsets) - 1
    deg = np.diff(row_offsets)
    out = []
    j = 0
    ms = mix64(np.uint64(seed))
    while len(out) < k:
        with np.errstate(over="ignore"):
            v = int(mix64(ms + np.uint64(j) * GOLDEN) % np.uint64(n))
        j += 1
        if deg[v] > 0:
            out.append(v)
    return np.array(out, np.int32)
