// Shared host/device helpers of the frontier searches.
#include <thread>
#include <vector>

#include <cstdio>

#include "frontier.h"

namespace mg {

__global__ void fill_i32(int *p, int v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void fill_u32(uint32_t *p, uint32_t v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void iota_i32(int *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (int)i;
}

int launch_fill(mg_graph *g, int *p, int v, int64_t n) {
    hipLaunchKernelGGL(fill_i32, dim3(grid_blocks(n, BLK)), dim3(BLK), 0, g->stream, p, v, n);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int launch_fill(mg_graph *g, uint32_t *p, uint32_t v, int64_t n) {
    hipLaunchKernelGGL(fill_u32, dim3(grid_blocks(n, BLK)), dim3(BLK), 0, g->stream, p, v, n);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

FrontierArgs make_frontier_args(mg_graph *g, int directed) {
    Workspace &ws = g->ws;
    FrontierArgs fa{};
    fa.ctl = ws.ctl;
    fa.off[0] = g->row_off;
    fa.adj[0] = g->col;
    fa.eid[0] = nullptr;
    fa.nviews = 1;
    if (directed == 0 && !(g->flags & MG_GRAPH_SYMMETRIC) && g->col_off) {
        fa.off[1] = g->col_off;
        fa.adj[1] = g->row_idx;
        fa.eid[1] = g->csc_eid;
        fa.nviews = 2;
    } else {
        fa.off[1] = fa.off[0];
        fa.adj[1] = fa.adj[0];
        fa.eid[1] = nullptr;
    }
    fa.off[2] = fa.adj[2] = fa.eid[2] = nullptr;
    if (g->flags & MG_GRAPH_SYMMETRIC) {
        fa.off[2] = g->row_off;
        fa.adj[2] = g->col;
    } else if (directed == 1 && g->col_off) {
        fa.off[2] = g->col_off;
        fa.adj[2] = g->row_idx;
        fa.eid[2] = g->csc_eid;
    }
    for (int i = 0; i < 3; ++i) {
        fa.bm[i] = ws.bm[i];
        fa.q[i] = ws.q[i];
    }
    fa.trace = ws.trace;
    fa.chunks = ws.chunks;
    fa.chunk_cap = ws.chunk_cap;
    fa.part = ws.part;
    fa.n = g->n;
    fa.nwords = ws.nwords;
    return fa;
}

// Static pull tasks: in-lists longer than `longer` cut into PULL_TASK-slot pieces.
static int build_pull_tasks(mg_graph *g, const int *d_in_off, int longer, ChunkRec **tasks, int *ntasks,
                            const int32_t **view) {
    if (*view == d_in_off) return MG_OK;
    const int n = g->n;
    std::vector<int32_t> off((size_t)n + 1);
    MG_HIP(copy_on(g->stream, off.data(), d_in_off, 4 * ((size_t)n + 1), hipMemcpyDeviceToHost));
    std::vector<ChunkRec> t;
    for (int v = 0; v < n; ++v) {
        if (off[v + 1] - off[v] <= longer) continue;
        for (int b = off[v]; b < off[v + 1]; b += PULL_TASK)
            t.push_back(ChunkRec{v, b, std::min(b + PULL_TASK, off[v + 1]), 2});
    }
    free_dev(*tasks);
    *tasks = nullptr;
    *view = nullptr;
    int rc = alloc_chunks(tasks, t.size());
    if (rc) return rc;
    if (!t.empty()) MG_HIP(copy_on(g->stream, *tasks, t.data(), sizeof(ChunkRec) * t.size(), hipMemcpyHostToDevice));
    *ntasks = (int)t.size();
    *view = d_in_off;
    return MG_OK;
}

int ensure_pull_tasks(mg_graph *g, const int *d_in_off) {
    Workspace &ws = g->ws;
    return build_pull_tasks(g, d_in_off, LANE_T, &ws.pull_tasks, &ws.pull_ntasks, &ws.pull_view);
}

int ensure_dense_tasks(mg_graph *g, const int *d_in_off) {
    Workspace &ws = g->ws;
    if (ws.dense_view == d_in_off) return MG_OK;
    const int n = g->n;
    std::vector<int32_t> off((size_t)n + 1);
    MG_HIP(copy_on(g->stream, off.data(), d_in_off, 4 * ((size_t)n + 1), hipMemcpyDeviceToHost));
    std::vector<ChunkRec> t;
    // pieces of the long lists
    for (int v = 0; v < n; ++v)
        if (off[v + 1] - off[v] > DENSE_BIG)
            for (int b = off[v]; b < off[v + 1]; b += DENSE_T)
                t.push_back(ChunkRec{v, b, std::min(b + DENSE_T, off[v + 1]), 0});
    // ranges of whole words over the other vertices (ranges without in-edges are dropped)
    int va = 0;
    int64_t edges = 0;
    for (int w0 = 0; w0 < n; w0 += 32) {
        const int w1 = std::min(w0 + 32, n);
        int64_t we = 0;
        for (int v = w0; v < w1; ++v) {
            const int d = off[v + 1] - off[v];
            if (d <= DENSE_BIG) we += d;
        }
        if (w0 > va && (edges + we > DENSE_T || w0 - va >= DENSE_V)) {
            if (edges) t.push_back(ChunkRec{-1, va, w0, 0});
            va = w0;
            edges = 0;
        }
        edges += we;
        if (edges > DENSE_RANGE_MAX) return set_error(MG_ERR_OVERFLOW, "dense range of %lld in-edges", (long long)edges);
    }
    if (edges) t.push_back(ChunkRec{-1, va, n, 0});
    free_dev(ws.dense_tasks);
    ws.dense_tasks = nullptr;
    ws.dense_view = nullptr;
    int rc = alloc_chunks(&ws.dense_tasks, t.size());
    if (rc) return rc;
    if (!t.empty()) MG_HIP(copy_on(g->stream, ws.dense_tasks, t.data(), sizeof(ChunkRec) * t.size(), hipMemcpyHostToDevice));
    ws.dense_ntasks = (int)t.size();
    ws.dense_view = d_in_off;
    return MG_OK;
}

// Reference TEPS count (bfs.cu:483-493) and reached-vertex count over a result
// array: sum of CSR out-degrees of vertices whose value differs from `unreached`.
__global__ __launch_bounds__(1024) void reached_stats(const int *vals, int unreached, const int *row_off, int n,
                                                      unsigned long long *out) {
    long long te = 0, nr = 0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        if (vals[v] != unreached) {
            te += row_off[v + 1] - row_off[v];
            nr++;
        }
    te = block_sum<1024>(te);
    nr = block_sum<1024>(nr);
    if (threadIdx.x == 0) {  // one atomic per workgroup (at most 256 workgroups)
        if (te) atomicAdd(&out[0], (unsigned long long)te);
        if (nr) atomicAdd(&out[1], (unsigned long long)nr);
    }
}

// Queues the TEPS count over `vals` on the graph's stream; the two counts land in the
// pinned ws.h_reach once the stream has drained.
static int queue_reached_stats(mg_graph *g, const int *vals, int unreached) {
    unsigned long long *d = g->ws.d_reach;
    MG_HIP(hipMemsetAsync(d, 0, 16, g->stream));
    const int blocks = std::min(256, std::max(1, (g->n + 1023) / 1024));
    hipLaunchKernelGGL(reached_stats, dim3(blocks), dim3(1024), 0, g->stream, vals, unreached, g->row_off, g->n, d);
    MG_HIP(hipGetLastError());
    MG_HIP(hipMemcpyAsync(g->ws.h_reach, d, 16, hipMemcpyDeviceToHost, g->stream));
    return MG_OK;
}

int print_trace(mg_graph *g, const char *what, const char *dir0, const char *dir1) {
    Workspace &ws = g->ws;
    MG_HIP(copy_on(g->stream, ws.h_ctl, ws.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    const int levels = ws.h_ctl->level;
    const int nl = std::min(levels, TRACE_CAP);
    std::vector<long long> t(2 * (size_t)std::max(nl, 1));
    if (nl) MG_HIP(copy_on(g->stream, t.data(), ws.trace, 16 * (size_t)nl, hipMemcpyDeviceToHost));
    for (int L = 0; L < nl; ++L) {
        const long long nf = t[2 * L] & ((1LL << 40) - 1);
        const int dir = (int)(t[2 * L] >> 40);
        std::fprintf(stderr, "%s level %d (%s): next frontier %lld vertices, %lld out-edges\n", what, L,
                     dir ? dir1 : dir0, nf, t[2 * L + 1]);
    }
    if (levels > nl) std::fprintf(stderr, "%s: %d more levels not recorded\n", what, levels - nl);
    return MG_OK;
}

// Records the closing event, fills *stats from the control block and copies
// `bytes` of the device result `result` to `out`.  `vals` (nullable) is the int
// result the TEPS count reads.
// out[v] = vals[perm[v]]: a result of the degree-ordered twin in the caller's numbering
__global__ void unpermute_vals(const int *perm, const int *vals, int n, int *out) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) out[v] = vals[perm[v]];
}

int launch_unpermute(const int *perm, const void *vals, int n, void *out, hipStream_t st) {
    hipLaunchKernelGGL(unpermute_vals, dim3(grid_blocks_cap(n, BLK, 4 * MAX_GRID)), dim3(BLK), 0, st, perm,
                       (const int *)vals, n, (int *)out);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int queue_finish(mg_graph *g, bool with_stats, const int *vals, int unreached, const void *result, size_t bytes) {
    if (g->perm && result && bytes == 4 * (size_t)g->n) {
        // a search on a twin: its per-vertex result (int32 or float) goes back to the caller's ids
        hipLaunchKernelGGL(unpermute_vals, dim3(grid_blocks_cap(g->n, BLK, 4 * MAX_GRID)), dim3(BLK), 0, g->stream,
                           g->perm, (const int *)result, g->n, g->unperm);
        MG_HIP(hipGetLastError());
    }
    MG_HIP(hipEventRecord(g->ev1, g->stream));  // the end of the timed search
    if (with_stats && vals) return queue_reached_stats(g, vals, unreached);
    return MG_OK;
}

int finish_search(mg_graph *g, mg_stats *stats, const int *vals, int unreached, void *out, const void *result,
                  size_t bytes, bool out_on_device, const CallTimer &timer, bool tail_queued) {
    Workspace &ws = g->ws;
    const bool reach = stats && vals;
    if (!tail_queued) {
        // the closing event, then the control block and the TEPS count behind it: one wait
        int rc = queue_finish(g, stats != nullptr, vals, unreached, result, bytes);
        if (rc) return rc;
        MG_HIP(hipMemcpyAsync(ws.h_ctl, ws.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, g->stream));
        MG_HIP(hipStreamSynchronize(g->stream));
    }  // else drive queued the same work and waited for it
    if (g->perm && result && bytes == 4 * (size_t)g->n) result = g->unperm;
    float ms = 0;
    MG_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
    if (ws.h_ctl->changed & (1 << 30))
        return set_error(MG_ERR_OVERFLOW, "edge-chunk queue overflow (capacity %lld)", (long long)ws.chunk_cap);
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->iterations = ws.h_ctl->level;
        stats->bottom_up_levels = ws.h_ctl->bu_levels;
        stats->edges_examined = ws.h_ctl->edges_examined;
        stats->device_ms = ms;
        if (reach) {
            stats->edges_traversed = (int64_t)ws.h_reach[0];
            stats->vertices_reached = (int64_t)ws.h_reach[1];
        }
    }
    int rc = deliver(g, out, result, bytes, out_on_device);
    if (rc) return rc;
    if (stats) stats->host_ms = timer.ms();
    return MG_OK;
}

}  // namespace mg
