// Device graph store and the shared C-ABI plumbing of libmapgraph_hip.so.
//
// mg_graph_create plays the part of CsrProblem::FromHostProblem
// (GASengine/csr_problem.cuh:401-621): it copies the host CSR (and, on request,
// builds the in-edge CSC view with the CSR edge id of every CSC slot, as
// util.cuh:55-126 does) into HBM.  Unlike the reference it allocates no O(m)
// frontier queues: the engine's frontiers are vertex queues and bitmaps of O(n)
// (the reference sizes three key/value queue pairs at max_queue_sizing * m,
// csr_problem.cuh:627-827), pooled per graph and reused across searches.
#include <climits>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "engine.h"

namespace mg {

static thread_local int t_device = 0;

int grid_blocks(int64_t work_items, int block) { return grid_blocks_cap(work_items, block, MAX_GRID); }

int grid_blocks_cap(int64_t work_items, int block, int cap) {
    int64_t b = (work_items + block - 1) / block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;  // kernels stride beyond this
    return (int)b;
}

bool alloc_refused() {
    static std::mutex mu;
    static std::string armed;
    static long long seen = 0;
    const char *e = getenv("MG_FAIL_ALLOC_AT");
    std::lock_guard<std::mutex> lock(mu);
    const std::string cur = e ? e : "";
    if (cur != armed) {
        armed = cur;
        seen = 0;
    }
    if (cur.empty()) return false;
    const long long k = atoll(cur.c_str());
    return k > 0 && ++seen == k;
}

template <class T>
static int dalloc(T **p, size_t count) {
    if (count == 0) count = 1;
    if (alloc_refused()) {
        *p = nullptr;
        return set_error(MG_ERR_OOM, "device allocation of %zu bytes refused (MG_FAIL_ALLOC_AT)", count * sizeof(T));
    }
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(MG_ERR_OOM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
    }
    return MG_OK;
}

template <class T>
static void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

static void free_workspace(Workspace &ws) {
    dfree(ws.vals);
    for (auto &p : ws.q) dfree(p);
    dfree(ws.hard);
    dfree(ws.visited);
    for (auto &p : ws.bm) dfree(p);
    dfree(ws.chunks);
    dfree(ws.part);
    dfree(ws.ctl);
    if (ws.h_ctl) (void)hipHostFree(ws.h_ctl);
    if (ws.h_reach) (void)hipHostFree(ws.h_reach);
    dfree(ws.d_reach);
    dfree(ws.rank);
    dfree(ws.contrib[0]);
    dfree(ws.contrib[1]);
    dfree(ws.partials);
    dfree(ws.pr_mid);
    dfree(ws.pr_chunks);
    dfree(ws.pr_partials);
    dfree(ws.pr_cpart);
    dfree(ws.pr_part);
    dfree(ws.pr_pact);
    dfree(ws.pr_bigv);
    dfree(ws.trace);
    dfree(ws.pull_tasks);
    dfree(ws.dense_tasks);
    ws = Workspace();
}

int ensure_workspace(mg_graph *g) {
    Workspace &ws = g->ws;
    if (ws.ctl) return MG_OK;
    const int n = g->n;
    ws.n = n;
    ws.nwords = ((n + 63) / 64) * 2;
    // records per level: TASK-edge pieces of lists longer than LANE_T (CHUNK of
    // lists longer than CHUNK_T in the tile path), over both push views
    ws.chunk_need = 2 * (g->m / TASK + g->m / (LANE_T + 1)) + 2 * (g->m / CHUNK + g->m / CHUNK_T) + 64;
    ws.chunk_cap = ws.chunk_need;
    int rc = 0;
    if ((rc = dalloc(&ws.vals, n)) || (rc = dalloc(&ws.q[0], n)) || (rc = dalloc(&ws.q[1], n)) ||
        (rc = dalloc(&ws.q[2], n)) || (rc = dalloc(&ws.hard, n)) || (rc = dalloc(&ws.visited, ws.nwords)) ||
        (rc = dalloc(&ws.bm[0], ws.nwords)) || (rc = dalloc(&ws.bm[1], ws.nwords)) ||
        (rc = dalloc(&ws.bm[2], ws.nwords)) || (rc = dalloc(&ws.chunks, ws.chunk_cap)) ||
        (rc = dalloc(&ws.part, 3 * BIG_GRID)) || (rc = dalloc(&ws.ctl, 1)) || (rc = dalloc(&ws.d_reach, 2)) ||
        (rc = dalloc(&ws.trace, 2 * (size_t)TRACE_CAP))) {
        free_workspace(ws);
        return rc;
    }
    hipError_t e = hipHostMalloc((void **)&ws.h_ctl, sizeof(Ctl), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&ws.h_reach, 2 * sizeof(unsigned long long), hipHostMallocDefault);
    if (e != hipSuccess) {
        free_workspace(ws);
        return set_error(MG_ERR_OOM, "hipHostMalloc failed: %s", hipGetErrorString(e));
    }
    return MG_OK;
}

int ensure_pr_workspace(mg_graph *g) {
    Workspace &ws = g->ws;
    if (ws.rank) return MG_OK;
    int rc = 0;
    const int n = g->n;
    if ((rc = dalloc(&ws.rank, n)) || (rc = dalloc(&ws.contrib[0], n)) || (rc = dalloc(&ws.contrib[1], n)) ||
        (rc = dalloc(&ws.partials, 8192))) {
        // all or nothing: ws.rank marks the PageRank state as present
        dfree(ws.rank);
        dfree(ws.contrib[0]);
        dfree(ws.contrib[1]);
        dfree(ws.partials);
        return rc;
    }
    return MG_OK;
}

int64_t queue_capacity(int64_t need, double max_queue_sizing) {
    if (!(max_queue_sizing > 0)) return need;  // 0, negative or NaN: the default
    const double c = (double)need * (max_queue_sizing / 1.5);
    return c < 64 ? 64 : c > 4e18 ? (int64_t)4e18 : (int64_t)c;
}

void invalidate_steps(mg_graph *g) {
    if (!g->steps) return;
    for (StepGraph &s : g->steps->sg) {
        if (s.exec) (void)hipGraphExecDestroy(s.exec);
        for (hipGraphExec_t &x : s.part) {
            if (x) (void)hipGraphExecDestroy(x);
            x = nullptr;
        }
        s = StepGraph();
    }
}

static int set_chunk_cap(mg_graph *g, int64_t cap) {
    Workspace &ws = g->ws;
    if (cap == ws.chunk_cap) return MG_OK;
    MG_HIP(hipStreamSynchronize(g->stream));  // no queued step may still use the old queue
    invalidate_steps(g);
    dfree(ws.chunks);
    ws.chunk_cap = 0;
    int rc = dalloc(&ws.chunks, (size_t)cap);
    if (rc) return rc;
    ws.chunk_cap = cap;
    return MG_OK;
}

int size_chunk_queue(mg_graph *g, const mg_options &o) {
    const Workspace &ws = g->ws;
    return set_chunk_cap(g, std::max(queue_capacity(ws.chunk_need, o.max_queue_sizing), ws.chunk_floor));
}

int grow_chunk_queue(mg_graph *g) {
    Workspace &ws = g->ws;
    ws.chunk_floor = 4 * ws.chunk_cap;
    return set_chunk_cap(g, ws.chunk_floor);
}

int alloc_i32(int32_t **p, size_t n) { return dalloc(p, n); }
int alloc_f32(float **p, size_t n) { return dalloc(p, n); }
int alloc_chunks(ChunkRec **p, size_t n) { return dalloc(p, n); }
int alloc_f64(double **p, size_t n) { return dalloc(p, n); }
int alloc_parts(LevelPart **p, size_t n) { return dalloc(p, n); }
void free_dev(void *p) {
    if (p) (void)hipFree(p);
}

bool graphs_enabled() {
    static const bool on = [] {
        const char *e = getenv("MG_NO_GRAPH");
        return !(e && *e && *e != '0');
    }();
    return on;
}

int begin_search(mg_graph *g, const mg_options *opt, mg_options *resolved) {
    clear_error();
    if (!g) return set_error(MG_ERR_INVALID_ARG, "null graph");
    if (opt) *resolved = *opt;
    else mg_options_default(resolved);
    if (resolved->iter_num < 0) return set_error(MG_ERR_INVALID_ARG, "iter_num must be >= 0");
    if (resolved->directed != 0 && resolved->directed != 1)
        return set_error(MG_ERR_INVALID_ARG, "directed must be 0 or 1");
    if (resolved->engine != MG_ENGINE_AUTO && resolved->engine != MG_ENGINE_GAS)
        return set_error(MG_ERR_INVALID_ARG, "unknown engine %d", resolved->engine);
    if (resolved->layout != 0 && resolved->layout != 1) return set_error(MG_ERR_INVALID_ARG, "layout must be 0 or 1");
    if (resolved->level_grid < 0 || resolved->level_grid > 65536)
        return set_error(MG_ERR_INVALID_ARG, "level_grid must be in [0, 65536]");
    if (resolved->directed == 0 && !(g->flags & (MG_GRAPH_BUILD_CSC | MG_GRAPH_SYMMETRIC)))
        return set_error(MG_ERR_UNSUPPORTED, "directed=0 needs a graph created with MG_GRAPH_BUILD_CSC or MG_GRAPH_SYMMETRIC");
    MG_HIP(hipSetDevice(g->device));
    return ensure_workspace(g);
}

// A graph object on `device` with its stream and events and no arrays yet (the
// caller fills them; mg_graph_destroy releases a partly built one).
static int graph_shell(int device, int32_t n, int64_t m, uint32_t flags, mg_graph **out) {
    *out = nullptr;
    mg_graph *g = new (std::nothrow) mg_graph();
    if (!g) return set_error(MG_ERR_OOM, "host allocation failed");
    g->steps = new (std::nothrow) mg_graph_steps();
    if (!g->steps) {
        delete g;
        return set_error(MG_ERR_OOM, "host allocation failed");
    }
    g->device = device;
    g->n = n;
    g->m = m;
    g->flags = flags;
    if (hipSetDevice(g->device) != hipSuccess) {
        mg_graph_destroy(g);
        return set_error(MG_ERR_DEVICE, "hipSetDevice failed");
    }
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&g->ev0) != hipSuccess || hipEventCreate(&g->ev1) != hipSuccess) {
        mg_graph_destroy(g);
        return set_error(MG_ERR_DEVICE, "stream/event creation failed");
    }
    *out = g;
    return MG_OK;
}

// Part p of a disjoint union: its offsets shifted by the edges before it, its column
// ids by the vertices before it.
__global__ void union_offsets(const int *ro, int n, int eoff, int *out) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) out[v] = ro[v] + eoff;
}
__global__ void union_columns(const int *col, int64_t m, int voff, int *out) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x)
        out[e] = col[e] + voff;
}

int deliver(mg_graph *g, void *dst, const void *src, size_t bytes, bool out_on_device) {
    if (!dst) return MG_OK;
    MG_HIP(hipMemcpyAsync(dst, src, bytes, out_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                          g->stream));
    MG_HIP(hipStreamSynchronize(g->stream));
    return MG_OK;
}

hipError_t dedicated_stream(int device, hipStream_t *st) {
    int cus = 0;
    hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) return e;
    std::vector<uint32_t> mask((size_t)std::max(1, (cus + 31) / 32), 0xffffffffu);
    return hipExtStreamCreateWithCUMask(st, (uint32_t)mask.size(), mask.data());
}

// The arrays and weight summary of `from` on the view `v` (borrowed).
static void borrow_arrays(mg_graph *v, const mg_graph *from) {
    v->arrays_borrowed = true;
    v->no_capture = true;
    v->row_off = from->row_off;
    v->col = from->col;
    v->w = from->w;
    v->col_off = from->col_off;
    v->row_idx = from->row_idx;
    v->csc_eid = from->csc_eid;
    v->min_weight = from->min_weight;
    v->max_weight = from->max_weight;
    v->mean_weight = from->mean_weight;
    v->perm = from->perm;
}

int graph_view_create(mg_graph *g, mg_graph **out) {
    *out = nullptr;
    mg_graph *v = nullptr;
    int rc = graph_shell(g->device, g->n, g->m, g->flags & ~MG_GRAPH_RELABEL, &v);
    if (rc) return rc;
    auto fail = [&](int code) {
        mg_graph_destroy(v);
        return code;
    };
    (void)hipStreamDestroy(v->stream);
    v->stream = nullptr;
    if (dedicated_stream(g->device, &v->stream) != hipSuccess) {
        v->stream = nullptr;
        return fail(set_error(MG_ERR_DEVICE, "dedicated stream creation failed"));
    }
    borrow_arrays(v, g);
    if (g->twin) {
        mg_graph *t = nullptr;
        if ((rc = graph_shell(g->device, g->n, g->m, g->twin->flags, &t))) return fail(rc);
        v->twin = t;  // released with the view from here on
        (void)hipStreamDestroy(t->stream);
        t->stream = v->stream;
        t->stream_borrowed = true;
        borrow_arrays(t, g->twin);
        if ((rc = dalloc(&t->unperm, (size_t)g->n))) return fail(rc);
        v->h_perm = g->h_perm;
        if ((rc = ensure_workspace(t))) return fail(rc);
    }
    if ((rc = ensure_workspace(v))) return fail(rc);
    *out = v;
    return MG_OK;
}

}  // namespace mg

using namespace mg;

extern "C" {

void mg_options_default(mg_options *o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->iter_num = INT_MAX;
    o->threshold = 10000;
    o->directed = 1;
    o->engine = MG_ENGINE_AUTO;
    o->pr_mode = MG_PR_GAS;
    o->pr_teleport = 0.15f;
    o->pr_tol = 0.0f;
    o->pr_active_eps = 0.01f;
    o->max_queue_sizing = 1.5;
    o->alpha = 14;
    o->beta = 24;
}

int mg_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int mg_device_init(int device) {
    clear_error();
    int c = mg_device_count();
    if (c == 0) return set_error(MG_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= c) return set_error(MG_ERR_INVALID_ARG, "device %d outside [0,%d)", device, c);
    MG_HIP(hipSetDevice(device));
    t_device = device;
    return MG_OK;
}

int mg_device_synchronize(void) {
    clear_error();
    MG_HIP(hipDeviceSynchronize());
    return MG_OK;
}

int mg_graph_create(const int32_t *row_offsets, const int32_t *col_idx, const int32_t *edge_values, int32_t n,
                    int64_t m, uint32_t flags, mg_graph **out) {
    clear_error();
    if (!out) return set_error(MG_ERR_INVALID_ARG, "mg_graph_create: out is null");
    *out = nullptr;
    if (n < 0 || m < 0 || m > INT32_MAX || !row_offsets || (m > 0 && !col_idx))
        return set_error(MG_ERR_INVALID_ARG, "mg_graph_create: bad sizes or null arrays (n=%d, m=%lld)", n,
                         (long long)m);
    if ((flags & MG_GRAPH_WEIGHTED) && m > 0 && !edge_values)
        return set_error(MG_ERR_INVALID_ARG, "mg_graph_create: MG_GRAPH_WEIGHTED without edge_values");
    if (flags & ~(MG_GRAPH_WEIGHTED | MG_GRAPH_BUILD_CSC | MG_GRAPH_SYMMETRIC | MG_GRAPH_RELABEL))
        return set_error(MG_ERR_INVALID_ARG, "mg_graph_create: unknown flags 0x%x", flags);
    if (row_offsets[0] != 0 || row_offsets[n] != m)
        return set_error(MG_ERR_INVALID_ARG, "mg_graph_create: row_offsets must start at 0 and end at m");
    for (int32_t v = 0; v < n; ++v)
        if (row_offsets[v + 1] < row_offsets[v])
            return set_error(MG_ERR_INVALID_ARG, "mg_graph_create: row_offsets decrease at %d", v);
    for (int64_t e = 0; e < m; ++e)
        if ((uint32_t)col_idx[e] >= (uint32_t)n)
            return set_error(MG_ERR_INVALID_ARG, "mg_graph_create: col_idx[%lld]=%d outside [0,%d)", (long long)e,
                             col_idx[e], n);
    int ndev = mg_device_count();
    if (ndev == 0) return set_error(MG_ERR_NO_DEVICE, "no HIP device visible");

    mg_graph *g = nullptr;
    int rc = graph_shell(t_device, n, m, flags, &g);
    if (rc) return rc;
    auto fail = [&](int code) {
        mg_graph_destroy(g);
        return code;
    };
    if ((rc = dalloc(&g->row_off, (size_t)n + 1)) || (rc = dalloc(&g->col, (size_t)m))) return fail(rc);
    g->topo_bytes = 4 * ((int64_t)n + 1 + m);
    if (copy_on(g->stream, g->row_off, row_offsets, 4 * ((size_t)n + 1), hipMemcpyHostToDevice) != hipSuccess ||
        copy_on(g->stream, g->col, col_idx, 4 * (size_t)m, hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_error(MG_ERR_DEVICE, "host-to-device copy of the CSR failed"));
    if (flags & MG_GRAPH_WEIGHTED) {
        if ((rc = dalloc(&g->w, (size_t)m))) return fail(rc);
        if (copy_on(g->stream, g->w, edge_values, 4 * (size_t)m, hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_error(MG_ERR_DEVICE, "host-to-device copy of the edge values failed"));
        int32_t mn = INT32_MAX, mx = INT32_MIN;
        double sum = 0;
        for (int64_t e = 0; e < m; ++e) {
            mn = edge_values[e] < mn ? edge_values[e] : mn;
            mx = edge_values[e] > mx ? edge_values[e] : mx;
            sum += edge_values[e];
        }
        g->min_weight = m ? mn : 0;
        g->max_weight = m ? mx : 0;
        g->mean_weight = m ? sum / (double)m : 0;
        g->topo_bytes += 4 * m;
    }
    if (flags & MG_GRAPH_BUILD_CSC) {
        std::vector<int32_t> co((size_t)n + 1), ri((size_t)m), ce((size_t)m);
        if ((rc = mg_csc_from_csr(n, m, row_offsets, col_idx, co.data(), ri.data(), ce.data()))) return fail(rc);
        if ((rc = dalloc(&g->col_off, (size_t)n + 1)) || (rc = dalloc(&g->row_idx, (size_t)m)) ||
            (rc = dalloc(&g->csc_eid, (size_t)m)))
            return fail(rc);
        if (copy_on(g->stream, g->col_off, co.data(), 4 * ((size_t)n + 1), hipMemcpyHostToDevice) != hipSuccess ||
            copy_on(g->stream, g->row_idx, ri.data(), 4 * (size_t)m, hipMemcpyHostToDevice) != hipSuccess ||
            copy_on(g->stream, g->csc_eid, ce.data(), 4 * (size_t)m, hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_error(MG_ERR_DEVICE, "host-to-device copy of the CSC failed"));
        g->topo_bytes += 4 * ((int64_t)n + 1 + 2 * m);
    }
    if (flags & MG_GRAPH_RELABEL) {
        // the degree-ordered twin: renamed on the host, then created like any graph
        const bool wt = (flags & MG_GRAPH_WEIGHTED) != 0;
        std::vector<int32_t> perm((size_t)n), ro2((size_t)n + 1), col2((size_t)m), w2(wt ? (size_t)m : 0);
        if ((rc = mg_csr_relabel_degree(n, m, row_offsets, col_idx, wt ? edge_values : nullptr, perm.data(),
                                        ro2.data(), col2.data(), wt ? w2.data() : nullptr)))
            return fail(rc);
        mg_graph *t = nullptr;
        if ((rc = mg_graph_create(ro2.data(), col2.data(), wt ? w2.data() : nullptr, n, m, flags & ~MG_GRAPH_RELABEL,
                                  &t)))
            return fail(rc);
        g->twin = t;
        // the twin is searched only under this graph's lock: it runs on this graph's
        // stream (each stream holds a hardware queue while it exists)
        if (hipStreamSynchronize(t->stream) != hipSuccess)
            return fail(set_error(MG_ERR_DEVICE, "twin creation did not complete"));
        (void)hipStreamDestroy(t->stream);
        t->stream = g->stream;
        t->stream_borrowed = true;
        if ((rc = dalloc(&t->perm, (size_t)n)) || (rc = dalloc(&t->unperm, (size_t)n))) return fail(rc);
        if (copy_on(t->stream, t->perm, perm.data(), 4 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_error(MG_ERR_DEVICE, "host-to-device copy of the vertex renaming failed"));
        g->h_perm = std::move(perm);
    }
    if ((rc = ensure_workspace(g))) return fail(rc);
    *out = g;
    return MG_OK;
}

int mg_graph_union(int32_t ngraphs, mg_graph *const *parts, mg_graph **out, int32_t *vertex_offsets) {
    clear_error();
    if (!out) return set_error(MG_ERR_INVALID_ARG, "mg_graph_union: out is null");
    *out = nullptr;
    if (ngraphs < 1 || !parts) return set_error(MG_ERR_INVALID_ARG, "mg_graph_union: no graphs");
    int64_t N = 0, M = 0;
    uint32_t flags = MG_GRAPH_WEIGHTED | MG_GRAPH_SYMMETRIC;
    for (int i = 0; i < ngraphs; ++i) {
        const mg_graph *p = parts[i];
        if (!p) return set_error(MG_ERR_INVALID_ARG, "mg_graph_union: graph %d is null", i);
        if (p->device != parts[0]->device)
            return set_error(MG_ERR_INVALID_ARG, "mg_graph_union: graphs on different devices");
        N += p->n;
        M += p->m;
        flags &= p->flags;
    }
    if (N >= INT32_MAX || M > INT32_MAX)
        return set_error(MG_ERR_INVALID_ARG, "mg_graph_union: %lld vertices / %lld edges exceed int32", (long long)N,
                         (long long)M);
    mg_graph *g = nullptr;
    int rc = graph_shell(parts[0]->device, (int32_t)N, M, flags, &g);
    if (rc) return rc;
    auto fail = [&](int code) {
        mg_graph_destroy(g);
        return code;
    };
    if ((rc = dalloc(&g->row_off, (size_t)N + 1)) || (rc = dalloc(&g->col, (size_t)M))) return fail(rc);
    if ((flags & MG_GRAPH_WEIGHTED) && (rc = dalloc(&g->w, (size_t)M))) return fail(rc);
    g->topo_bytes = 4 * (N + 1 + M) + ((flags & MG_GRAPH_WEIGHTED) ? 4 * M : 0);
    int64_t voff = 0, eoff = 0;
    double wsum = 0;
    g->min_weight = INT32_MAX;
    g->max_weight = INT32_MIN;
    for (int i = 0; i < ngraphs; ++i) {
        const mg_graph *p = parts[i];
        if (vertex_offsets) vertex_offsets[i] = (int32_t)voff;
        // the part's own stream is idle past this point: its arrays are read on g's stream
        MG_HIP(hipStreamSynchronize(p->stream));
        if (p->n > 0)
            hipLaunchKernelGGL(union_offsets, dim3(grid_blocks(p->n, BLK)), dim3(BLK), 0, g->stream, p->row_off, p->n,
                               (int)eoff, g->row_off + voff);
        if (p->m > 0) {
            hipLaunchKernelGGL(union_columns, dim3(grid_blocks(p->m, BLK)), dim3(BLK), 0, g->stream, p->col, p->m,
                               (int)voff, g->col + eoff);
            if (flags & MG_GRAPH_WEIGHTED)
                MG_HIP(hipMemcpyAsync(g->w + eoff, p->w, 4 * (size_t)p->m, hipMemcpyDeviceToDevice, g->stream));
        }
        if (p->m > 0) {
            g->min_weight = std::min(g->min_weight, p->min_weight);
            g->max_weight = std::max(g->max_weight, p->max_weight);
            wsum += p->mean_weight * (double)p->m;
        }
        voff += p->n;
        eoff += p->m;
    }
    if (vertex_offsets) vertex_offsets[ngraphs] = (int32_t)voff;
    const int32_t last = (int32_t)M;
    MG_HIP(hipMemcpyAsync(g->row_off + N, &last, 4, hipMemcpyHostToDevice, g->stream));
    MG_HIP(hipGetLastError());
    MG_HIP(hipStreamSynchronize(g->stream));
    if (M == 0) g->min_weight = g->max_weight = 0;
    g->mean_weight = M ? wsum / (double)M : 0;
    if ((rc = ensure_workspace(g))) return fail(rc);
    *out = g;
    return MG_OK;
}

int mg_graph_destroy(mg_graph *g) {
    clear_error();
    if (!g) return MG_OK;
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->steps) {
        gas_release(g);
        batch_release(g);
        ppr_release(g);
        for (auto &s : g->steps->sg) {
            if (s.exec) (void)hipGraphExecDestroy(s.exec);
            for (hipGraphExec_t x : s.part)
                if (x) (void)hipGraphExecDestroy(x);
        }
        delete g->steps;
        g->steps = nullptr;
    }
    if (g->twin) mg_graph_destroy(g->twin);
    g->twin = nullptr;
    dfree(g->unperm);
    free_workspace(g->ws);
    if (!g->arrays_borrowed) {
        dfree(g->perm);
        dfree(g->row_off);
        dfree(g->col);
        dfree(g->w);
        dfree(g->col_off);
        dfree(g->row_idx);
        dfree(g->csc_eid);
    }
    if (g->ev0) (void)hipEventDestroy(g->ev0);
    if (g->ev1) (void)hipEventDestroy(g->ev1);
    if (g->stream && !g->stream_borrowed) (void)hipStreamDestroy(g->stream);
    delete g;
    return MG_OK;
}

int mg_graph_info(const mg_graph *g, int32_t *n, int64_t *m, uint32_t *flags) {
    clear_error();
    if (!g) return set_error(MG_ERR_INVALID_ARG, "null graph");
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (flags) *flags = g->flags;
    return MG_OK;
}

int mg_graph_device_bytes(const mg_graph *g, int64_t *bytes) {
    clear_error();
    if (!g || !bytes) return set_error(MG_ERR_INVALID_ARG, "null argument");
    const Workspace &ws = g->ws;
    int64_t b = g->topo_bytes;
    if (ws.ctl) b += 4LL * ws.n * 5 + 4LL * ws.nwords * 4 + (int64_t)sizeof(ChunkRec) * ws.chunk_cap;
    if (ws.rank) b += 12LL * ws.n;
    if (g->steps)
        for (const BatchDev *p : g->steps->bdev)
            if (p) b += batch_dev_size(p);
    if (g->steps && g->steps->pdev) b += ppr_dev_size(g->steps->pdev);
    if (g->perm) b += 8LL * g->n;
    if (g->twin) {
        int64_t tb = 0;
        mg_graph_device_bytes(g->twin, &tb);
        b += tb;
    }
    *bytes = b;
    return MG_OK;
}

int mg_graph_get_csc(const mg_graph *g, int32_t *col_offsets, int32_t *row_idx, int32_t *csr_edge_id) {
    clear_error();
    if (!g) return set_error(MG_ERR_INVALID_ARG, "null graph");
    if (!g->col_off) return set_error(MG_ERR_UNSUPPORTED, "graph has no CSC view (MG_GRAPH_BUILD_CSC)");
    MG_HIP(hipSetDevice(g->device));
    if (col_offsets) MG_HIP(copy_on(g->stream, col_offsets, g->col_off, 4 * ((size_t)g->n + 1), hipMemcpyDeviceToHost));
    if (row_idx) MG_HIP(copy_on(g->stream, row_idx, g->row_idx, 4 * (size_t)g->m, hipMemcpyDeviceToHost));
    if (csr_edge_id) MG_HIP(copy_on(g->stream, csr_edge_id, g->csc_eid, 4 * (size_t)g->m, hipMemcpyDeviceToHost));
    return MG_OK;
}

}  // extern "C"
