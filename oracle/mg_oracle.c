/*
 * mg_oracle.c -- CPU restatement of MapGraph's reference algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke()
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load it.
 * The product (libmapgraph_hip.so) never links or calls it.
 *
 * Each function cites the reference code (MapGraph 0.3.3 tree) it follows.  The
 * reference CPU oracles live inside .cu drivers that need CUDA headers, so they are
 * restated here in plain C over CSR arrays instead of compiled from the tree.
 *
 * Pinning (see tests/test_oracle.py): BFS/SSSP/CC are checked against the known
 * answers of the reference fixtures (smallRegressionGraphs: small, trivialGraph,
 * chesapeake), against scipy's
 * csgraph routines (breadth-first order, Dijkstra, connected components) on random
 * graphs, and against brute force on tiny graphs; the two SSSP variants are checked
 * against each other.  PageRank variants are pinned by closed forms (star, 2-cycle,
 * chain), by a float64 power iteration with numpy, and by hand-derived known answers
 * of the program's activation rule and of the Gauss-Seidel sweep order.
 *
 * The Jacobi PageRank loops run their per-vertex sums in parallel (OpenMP) over
 * destination vertices: each vertex's sum stays sequential in CSC order and the L1
 * change is summed in vertex order afterwards, so the output does not depend on the
 * thread count.
 */
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------ BFS
 * CPUBFS, Algorithms/BFS/bfs.cu:56-143: FIFO queue BFS; depth[src] = 0, unreached
 * = -1; when directed == 0 the in-edges (CSC: column_offsets/row_indices) are
 * walked too (bfs.cu:115-134).  Returns the search depth (levels). */
ORACLE_API int oracle_bfs(int directed, int n, const int *row_offsets, const int *col_idx,
                          const int *col_offsets, const int *row_idx, int src, int *depth) {
    for (int i = 0; i < n; i++) depth[i] = -1;
    if (src < 0 || src >= n) return -1;
    int *queue = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int head = 0, tail = 0, search_depth = 0;
    depth[src] = 0;
    queue[tail++] = src;
    while (head < tail) {
        int u = queue[head++];
        int nd = depth[u] + 1;
        for (int e = row_offsets[u]; e < row_offsets[u + 1]; e++) {
            int v = col_idx[e];
            if (depth[v] == -1) {
                depth[v] = nd;
                if (search_depth < nd) search_depth = nd;
                queue[tail++] = v;
            }
        }
        if (directed == 0) {
            for (int e = col_offsets[u]; e < col_offsets[u + 1]; e++) {
                int v = row_idx[e];
                if (depth[v] == -1) {
                    depth[v] = nd;
                    if (search_depth < nd) search_depth = nd;
                    queue[tail++] = v;
                }
            }
        }
    }
    free(queue);
    return search_depth + 1;
}

/* ------------------------------------------------------------------ SSSP
 * CPUSSSP, Algorithms/SSSP/sssp.cu:127-199: O(n^2) Dijkstra -- repeatedly pick the
 * unvisited vertex of smallest distance (linear scan, ties to the lowest id),
 * relax its out-edges with 64-bit sums (sssp.cu:173-178).  `unreached` is the
 * initial distance: the reference CPU uses 100000000 (sssp.cu:138-139), the GPU
 * program INT_MAX (sssp.h:43); the engine's contract is INT_MAX.  The selection
 * bound is INT_MAX as in sssp.cu:151.  For directed == 0 the reverse pass reads the
 * weight through the CSR edge id of the CSC slot (csc_edge_id), not the CSC
 * position the reference uses (sssp.cu:181-191). */
ORACLE_API void oracle_sssp_quadratic(int directed, int n, const int *row_offsets, const int *col_idx,
                                      const int *weights, const int *col_offsets, const int *row_idx,
                                      const int *csc_edge_id, int src, int unreached, int *dist) {
    char *visited = (char *)calloc((size_t)(n > 0 ? n : 1), 1);
    for (int i = 0; i < n; i++) dist[i] = unreached;
    if (src >= 0 && src < n) dist[src] = 0;
    for (;;) {
        int u = -1;
        int sd = INT_MAX;
        for (int i = 0; i < n; i++)
            if (!visited[i] && dist[i] < sd) {
                sd = dist[i];
                u = i;
            }
        if (u == -1) break;
        visited[u] = 1;
        for (int j = row_offsets[u]; j < row_offsets[u + 1]; ++j) {
            int v = col_idx[j];
            long long len = (long long)dist[u] + weights[j];
            if (len < dist[v]) dist[v] = (int)len;
        }
        if (!directed) {
            for (int j = col_offsets[u]; j < col_offsets[u + 1]; ++j) {
                int v = row_idx[j];
                long long len = (long long)dist[u] + weights[csc_edge_id[j]];
                if (len < dist[v]) dist[v] = (int)len;
            }
        }
    }
    free(visited);
}

/* Binary-heap Dijkstra with the same output as oracle_sssp_quadratic (integer
 * shortest distances are unique), used where the O(n^2) scan is too slow.  Lazy
 * deletion; 64-bit path sums, so no overflow below INT_MAX. */
typedef struct {
    long long d;
    int v;
} heap_item;

static void heap_push(heap_item *h, long long *sz, heap_item x) {
    long long i = (*sz)++;
    while (i > 0) {
        long long p = (i - 1) / 2;
        if (h[p].d < x.d || (h[p].d == x.d && h[p].v <= x.v)) break;
        h[i] = h[p];
        i = p;
    }
    h[i] = x;
}

static heap_item heap_pop(heap_item *h, long long *sz) {
    heap_item top = h[0], x = h[--(*sz)];
    long long i = 0;
    for (;;) {
        long long l = 2 * i + 1, r = l + 1, m = i;
        heap_item best = x;
        if (l < *sz && (h[l].d < best.d || (h[l].d == best.d && h[l].v < best.v))) {
            m = l;
            best = h[l];
        }
        if (r < *sz && (h[r].d < best.d || (h[r].d == best.d && h[r].v < best.v))) {
            m = r;
            best = h[r];
        }
        if (m == i) break;
        h[i] = h[m];
        i = m;
    }
    h[i] = x;
    return top;
}

ORACLE_API void oracle_sssp_heap(int directed, int n, const int *row_offsets, const int *col_idx,
                                 const int *weights, const int *col_offsets, const int *row_idx,
                                 const int *csc_edge_id, int src, int unreached, int *dist) {
    long long m = n > 0 ? row_offsets[n] : 0;
    long long cap = m + (directed ? 0 : m) + 2;
    heap_item *h = (heap_item *)malloc(sizeof(heap_item) * (size_t)cap);
    char *done = (char *)calloc((size_t)(n > 0 ? n : 1), 1);
    long long sz = 0;
    for (int i = 0; i < n; i++) dist[i] = unreached;
    if (src >= 0 && src < n) {
        dist[src] = 0;
        heap_item s = {0, src};
        heap_push(h, &sz, s);
    }
    while (sz > 0) {
        heap_item it = heap_pop(h, &sz);
        int u = it.v;
        if (done[u] || it.d != dist[u]) continue;
        done[u] = 1;
        for (int j = row_offsets[u]; j < row_offsets[u + 1]; ++j) {
            int v = col_idx[j];
            long long len = (long long)dist[u] + weights[j];
            if (len < dist[v]) {
                dist[v] = (int)len;
                heap_item x = {len, v};
                heap_push(h, &sz, x);
            }
        }
        if (!directed)
            for (int j = col_offsets[u]; j < col_offsets[u + 1]; ++j) {
                int v = row_idx[j];
                long long len = (long long)dist[u] + weights[csc_edge_id[j]];
                if (len < dist[v]) {
                    dist[v] = (int)len;
                    heap_item x = {len, v};
                    heap_push(h, &sz, x);
                }
            }
    }
    free(h);
    free(done);
}

/* ------------------------------------------------------------------ CC
 * CPUCC, Algorithms/CC/cc.cu:132-199: label[v] = v, then Gauss-Seidel sweeps --
 * forward over out-edges (CSR), reverse over in-edges (CSC) -- taking the minimum
 * neighbour label until no label changes.  Returns the number of sweeps. */
ORACLE_API int oracle_cc(int n, const int *row_offsets, const int *col_idx, const int *col_offsets,
                         const int *row_idx, int *label) {
    for (int i = 0; i < n; i++) label[i] = i;
    int changed = 1, sweeps = 0;
    while (changed) {
        changed = 0;
        sweeps++;
        for (int v = 0; v < n; v++) {
            int minnb = INT_MAX;
            for (int j = row_offsets[v]; j < row_offsets[v + 1]; j++) {
                int nb = col_idx[j];
                if (label[nb] < minnb) minnb = label[nb];
            }
            if (minnb < label[v]) {
                label[v] = minnb;
                changed = 1;
            }
        }
        for (int v = 0; v < n; v++) {
            int minnb = INT_MAX;
            for (int j = col_offsets[v]; j < col_offsets[v + 1]; j++) {
                int nb = row_idx[j];
                if (label[nb] < minnb) minnb = label[nb];
            }
            if (minnb < label[v]) {
                label[v] = minnb;
                changed = 1;
            }
        }
    }
    return sweeps;
}

/* Same labels as oracle_cc (min vertex id of each weakly connected component) by
 * union-find over the CSR edges: each union hangs the larger root under the smaller,
 * so a root is always its component's smallest id.  O(m alpha(n)) instead of the
 * sweep count times m; used as the checker on the largest graphs and pinned against
 * oracle_cc and scipy on smaller ones. */
static int uf_find(int *parent, int v) {
    while (parent[v] != v) {
        parent[v] = parent[parent[v]];
        v = parent[v];
    }
    return v;
}

ORACLE_API void oracle_cc_union_find(int n, const int *row_offsets, const int *col_idx, int *label) {
    for (int i = 0; i < n; i++) label[i] = i;
    for (int u = 0; u < n; u++)
        for (int j = row_offsets[u]; j < row_offsets[u + 1]; j++) {
            int a = uf_find(label, u), b = uf_find(label, col_idx[j]);
            if (a < b) label[b] = a;
            else if (b < a) label[a] = b;
        }
    for (int v = 0; v < n; v++) label[v] = uf_find(label, v);
}

/* ------------------------------------------------------------------ PageRank
 * Out-degree as CPUPR computes it: adjacent_difference of row_offsets
 * (PageRank.cu:121-125; PageRank.h:147-153 on the device). */
static int *out_degrees(int n, const int *row_offsets) {
    int *deg = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int v = 0; v < n; v++) deg[v] = row_offsets[v + 1] - row_offsets[v];
    return deg;
}

/* CPUPR, Algorithms/PageRank/PageRank.cu:108-163: Gauss-Seidel in vertex order,
 * float accumulation in in-edge (CSC) order, r = 0.15 + 0.85 * sum r[u]/outdeg[u],
 * repeated while any vertex moved by >= 0.01.  Returns the sweep count. */
ORACLE_API int oracle_pr_gauss_seidel(int n, const int *row_offsets, const int *col_offsets,
                                      const int *row_idx, int max_sweeps, float *rank) {
    int *deg = out_degrees(n, row_offsets);
    for (int i = 0; i < n; i++) rank[i] = 0.15f;
    int changed = 1, it = 0;
    while (changed && it < max_sweeps) {
        changed = 0;
        for (int v = 0; v < n; v++) {
            float s = 0.0f;
            for (int j = col_offsets[v]; j < col_offsets[v + 1]; ++j) {
                int nb = row_idx[j];
                s += rank[nb] / (float)deg[nb];
            }
            /* sumnb = 0.15 + 0.85 * sumnb: double constants, float store */
            s = (float)(0.15 + 0.85 * (double)s);
            if ((double)fabsf(s - rank[v]) >= 0.01) changed = 1; /* float difference, double compare */
            rank[v] = s;
        }
        it++;
    }
    free(deg);
    return it;
}

/* The GPU PageRank program as the enactor runs it (PageRank.h:192-283 driven by
 * enactor_vertex_centric.cuh:2526-2872):
 *   frontier_0 = all vertices, rank = 0.15 (PageRank.h:115-116, 130-139);
 *   each iteration: gather over in-edges of frontier vertices, gather_edge =
 *   rank[u] / (float)outdeg[u] (PageRank.h:233-243), summed with '+' (zero
 *   in-degree gives INIT_VALUE = 0); apply new = 0.15f + (1.0f - 0.15f) * g,
 *   changed = |old - new| >= 0.01f (PageRank.h:257-283), all applies after all
 *   gathers (Jacobi); expand: next frontier = out-neighbours of changed vertices,
 *   deduplicated (allow_duplicates = false, expand_atomic_flag / copy_if);
 *   stop when the frontier is empty or after iter_num iterations.
 * `teleport` and `eps` generalise the constants 0.15f and 0.01f; the damping is
 * derived as (1.0f - teleport) in float exactly as PageRank.h:269 writes it, so
 * teleport = 0.15f reproduces the reference constants.  Sums run sequentially in
 * CSC order in float.
 * Returns the iteration count. */
ORACLE_API int oracle_pr_gas(int n, const int *row_offsets, const int *col_idx, const int *col_offsets,
                             const int *row_idx, int iter_num, float teleport, float eps, float *rank) {
    int *deg = out_degrees(n, row_offsets);
    char *active = (char *)malloc((size_t)(n > 0 ? n : 1));
    char *changed = (char *)malloc((size_t)(n > 0 ? n : 1));
    float *g = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    const float base = teleport, damping = 1.0f - teleport;
    for (int i = 0; i < n; i++) {
        rank[i] = base;
        active[i] = 1;
    }
    int it = 0, nactive = n;
    while (it < iter_num && nactive > 0) {
        for (int v = 0; v < n; v++) {
            if (!active[v]) continue;
            float s = 0.0f;
            for (int j = col_offsets[v]; j < col_offsets[v + 1]; ++j) {
                int u = row_idx[j];
                s = s + rank[u] / (float)deg[u];
            }
            g[v] = s;
        }
        for (int v = 0; v < n; v++) {
            changed[v] = 0;
            if (!active[v]) continue;
            float nv = base + damping * g[v];
            changed[v] = !(fabsf(rank[v] - nv) < eps);
            rank[v] = nv;
        }
        nactive = 0;
        memset(active, 0, (size_t)(n > 0 ? n : 1));
        for (int u = 0; u < n; u++)
            if (changed[u])
                for (int e = row_offsets[u]; e < row_offsets[u + 1]; ++e) active[col_idx[e]] = 1;
        for (int v = 0; v < n; v++) nactive += active[v];
        it++;
    }
    free(deg);
    free(active);
    free(changed);
    free(g);
    return it;
}

/* All-vertex Jacobi PageRank in float (every vertex recomputed each iteration from
 * the previous iteration's ranks) in the reference scale and formula (r0 = t,
 * r = t + (1.0f - t) * sum rank[u]/outdeg[u] with t = teleport, dangling mass
 * dropped, PageRank.h:233-283).
 * Stops after iters iterations or when sum |r_new - r_old| < tol * n.  Sums run
 * sequentially in CSC order.  Returns the iteration count; *l1 gets the last L1
 * change. */
ORACLE_API int oracle_pr_jacobi_f32(int n, const int *row_offsets, const int *col_offsets, const int *row_idx,
                                    int iters, float teleport, double tol, float *rank, double *l1) {
    int *deg = out_degrees(n, row_offsets);
    float *contrib = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    double *dv = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    const float base = teleport, damping = 1.0f - teleport;
    for (int i = 0; i < n; i++) rank[i] = base;
    int it = 0;
    double delta = 0;
    while (it < iters) {
#pragma omp parallel for schedule(static)
        for (int u = 0; u < n; u++) contrib[u] = rank[u] / (float)deg[u];
#pragma omp parallel for schedule(dynamic, 4096)
        for (int v = 0; v < n; v++) {
            float s = 0.0f;
            for (int j = col_offsets[v]; j < col_offsets[v + 1]; ++j) s = s + contrib[row_idx[j]];
            float nv = base + damping * s;
            dv[v] = fabs((double)nv - (double)rank[v]);
            rank[v] = nv;
        }
        delta = 0;
        for (int v = 0; v < n; v++) delta += dv[v];
        it++;
        if (delta < tol * (double)n) break;
    }
    if (l1) *l1 = delta;
    free(deg);
    free(contrib);
    free(dv);
    return it;
}

/* Same iteration in float64 (the exact-arithmetic reference for FP32 error). */
ORACLE_API int oracle_pr_jacobi_f64(int n, const int *row_offsets, const int *col_offsets, const int *row_idx,
                                    int iters, double teleport, double tol, double *rank, double *l1) {
    int *deg = out_degrees(n, row_offsets);
    double *contrib = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double *dv = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    const double base = teleport, damping = 1.0 - teleport;
    for (int i = 0; i < n; i++) rank[i] = base;
    int it = 0;
    double delta = 0;
    while (it < iters) {
#pragma omp parallel for schedule(static)
        for (int u = 0; u < n; u++) contrib[u] = deg[u] ? rank[u] / (double)deg[u] : 0.0;
#pragma omp parallel for schedule(dynamic, 4096)
        for (int v = 0; v < n; v++) {
            double s = 0.0;
            for (int j = col_offsets[v]; j < col_offsets[v + 1]; ++j) s += contrib[row_idx[j]];
            double nv = base + damping * s;
            dv[v] = fabs(nv - rank[v]);
            rank[v] = nv;
        }
        delta = 0;
        for (int v = 0; v < n; v++) delta += dv[v];
        it++;
        if (delta < tol * (double)n) break;
    }
    if (l1) *l1 = delta;
    free(deg);
    free(contrib);
    free(dv);
    return it;
}

/* ------------------------------------------------------------------ TEPS
 * Traversed-edge count, Algorithms/BFS/bfs.cu:483-493: sum of out-degrees of the
 * vertices whose label differs from `unreached` (-1 for BFS). */
ORACLE_API long long oracle_teps_edges(int n, const int *row_offsets, const int *label, int unreached) {
    long long te = 0;
    for (int i = 0; i < n; i++)
        if (label[i] != unreached) te += row_offsets[i + 1] - row_offsets[i];
    return te;
}
