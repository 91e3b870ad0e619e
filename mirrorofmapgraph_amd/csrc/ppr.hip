// Batched personalized PageRank: up to 64 sources (slots) iterate together.
//
// Not in the reference, whose PageRank (Algorithms/PageRank/PageRank.h:233-283) has
// one global teleport term; this is its personalized form, with the reference's
// conventions: contribution r[v] / outdeg(v) gathered over in-edges, rank-sink
// (dangling) mass dropped, teleport t applied as written there (t + (1 - t) * sum):
//   x_0 = e_s,   x_{i+1}[u] = (1 - t) * sum_{v -> u} x_i[v] / outdeg(v)  (+ t when u = s).
// Ranks are vertex-major, X[v * 64 + slot] (256 bytes per vertex), so one wave
// gathers one in-edge for all slots (lane = slot).  Slots are handed out lowest
// first, so while at most 32 are in use only the first 128-byte line of a row is read
// and written.
// Vertices are covered by static tasks: runs of consecutive vertices with at most
// PPR_T in-edges in total (one wave walks them in order), and PPR_T-edge pieces
// of longer in-lists whose partial sums are added in piece order.  Every slot's
// arithmetic is thus fixed by the graph alone: a query gives the same bits alone
// or beside any other queries.
#include "batch.h"

#include <vector>

namespace mg {

constexpr int PPR_T = 1024;    // in-edges per task
constexpr int PPR_RUN = 64;    // vertices per run task

struct PprArgs {
    const int *in_off, *in_adj;  // in-edges (CSR of a symmetric graph, or the CSC)
    const int *out_off;          // out-degrees
    const float *X;              // current ranks, n x PPR_SLOTS
    float *Xn;                   // next ranks
    float *C;                    // contributions X / outdeg
    float *part;                 // piece partial sums, npart x PPR_SLOTS
    const ChunkRec *tasks;       // runs: {v0, -, -, count}; pieces: {v, beg, end, -(part + 1)}
    const int *big;              // per long list: vertex, first part, parts (3 ints)
    const int *src;              // per slot: source or -1
    int n, ntasks, nbig;
    float t;                     // teleport
};

__global__ __launch_bounds__(BLK) void ppr_contrib(PprArgs a) {
    const int lane = lane_id();
    const int wave = (blockIdx.x * BLK + (int)threadIdx.x) / WAVE, nwaves = gridDim.x * BLK / WAVE;
    const bool on = a.src[lane] >= 0;
    if (!__any(on)) return;
    for (int v = wave; v < a.n; v += nwaves) {
        const int d = a.out_off[v + 1] - a.out_off[v];
        if (on) a.C[(int64_t)v * PPR_SLOTS + lane] = d > 0 ? a.X[(int64_t)v * PPR_SLOTS + lane] / (float)d : 0.f;
    }
}

// Sum of C over in-edges [beg, end) in order, four loads in flight.
// Sum over in-edges [beg, end) of the lane's column, in edge order: full groups of
// PPR_U rows with every load in flight (the group test is wave-uniform, so the
// loads need no per-load branch), then the remainder one row at a time.
// `col` is the column the lane reads: its own slot, or, when no slot of the upper half
// of the row is in use (slots are handed out lowest first), the lower half's column
// of the same cache line, so a row costs one 128-byte line instead of two.
constexpr int PPR_U = 8;
__device__ __forceinline__ float gather_sum(const PprArgs &a, int beg, int end, int col, bool on) {
    float acc = 0.f;
    int e = beg;
    for (; e + PPR_U <= end; e += PPR_U) {
        int v[PPR_U];
        float c[PPR_U];
#pragma unroll
        for (int j = 0; j < PPR_U; ++j) v[j] = a.in_adj[e + j];
#pragma unroll
        for (int j = 0; j < PPR_U; ++j) c[j] = a.C[(int64_t)v[j] * PPR_SLOTS + col];
#pragma unroll
        for (int j = 0; j < PPR_U; ++j) acc += c[j];
    }
    for (; e < end; ++e) acc += a.C[(int64_t)a.in_adj[e] * PPR_SLOTS + col];
    return on ? acc : 0.f;
}

// The column lane `lane` reads: its own, unless the row's upper half is idle.
__device__ __forceinline__ int ppr_col(int lane, bool on) {
    const bool hi = __any(on && lane >= PPR_SLOTS / 2);
    return hi ? lane : (lane & (PPR_SLOTS / 2 - 1));
}

__device__ __forceinline__ float ppr_apply(float s, float t, bool is_src) {
    const float v = (1.0f - t) * s;
    return is_src ? t + v : v;
}

__global__ __launch_bounds__(BLK) void ppr_gather(PprArgs a) {
    const int lane = lane_id();
    const int wave = (blockIdx.x * BLK + (int)threadIdx.x) / WAVE, nwaves = gridDim.x * BLK / WAVE;
    const int src = a.src[lane];
    const bool on = src >= 0;
    if (!__any(on)) return;
    const int col = ppr_col(lane, on);
    for (int t = wave; t < a.ntasks; t += nwaves) {
        const ChunkRec r = a.tasks[t];
        if (r.pad > 0) {
            for (int u = r.v; u < r.v + r.pad; ++u) {
                const float s = gather_sum(a, a.in_off[u], a.in_off[u + 1], col, on);
                if (on) a.Xn[(int64_t)u * PPR_SLOTS + lane] = ppr_apply(s, a.t, u == src);  // idle columns unused
            }
        } else {
            const float s = gather_sum(a, r.beg, r.end, col, on);
            if (on) a.part[(int64_t)(-r.pad - 1) * PPR_SLOTS + lane] = s;
        }
    }
}

__global__ __launch_bounds__(BLK) void ppr_combine(PprArgs a) {
    const int lane = lane_id();
    const int wave = (blockIdx.x * BLK + (int)threadIdx.x) / WAVE, nwaves = gridDim.x * BLK / WAVE;
    const int src = a.src[lane];
    const bool on = src >= 0;
    for (int b = wave; b < a.nbig; b += nwaves) {
        const int u = a.big[3 * b], p0 = a.big[3 * b + 1], np = a.big[3 * b + 2];
        float s = 0.f;
        for (int p = p0; p < p0 + np; ++p) s += on ? a.part[(int64_t)p * PPR_SLOTS + lane] : 0.f;
        if (on) a.Xn[(int64_t)u * PPR_SLOTS + lane] = ppr_apply(s, a.t, u == src);
    }
}

// Column reset of the listed slots: x_0 = e_src.
__global__ void ppr_seed_cols(float *X, int n, SeedList sl) {
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, nt = (int64_t)gridDim.x * blockDim.x;
    for (int64_t u = tid; u < n; u += nt)
        for (int j = 0; j < sl.n; ++j) X[u * PPR_SLOTS + sl.slot[j]] = u == sl.src[j] ? 1.f : 0.f;
}

__global__ void ppr_set_src(int *src, SeedList sl, SlotList off) {
    const int i = threadIdx.x;
    if (i < off.n) src[off.slot[i]] = -1;
    __syncthreads();
    if (i < sl.n) src[sl.slot[i]] = sl.src[i];
}

// Rank columns of the listed slots -> their rows (64 x 64 LDS tile).
__global__ __launch_bounds__(BLK) void ppr_extract(const float *X, float *rows, int n, SlotList sl) {
    __shared__ float tile[64][65];
    const int tv = (n + 63) / 64;
    for (int t = blockIdx.x; t < tv; t += gridDim.x) {
        const int v0 = t * 64;
        for (int i = threadIdx.x; i < 64 * 64; i += BLK) {
            const int r = i / 64, cc = i % 64;
            tile[r][cc] = (v0 + r < n) ? X[(int64_t)(v0 + r) * PPR_SLOTS + cc] : 0.f;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < sl.n * 64; i += BLK) {
            const int j = i / 64, r = i % 64;
            if (v0 + r < n) rows[(int64_t)sl.slot[j] * n + v0 + r] = tile[r][sl.slot[j]];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- host side

void ppr_dev_free(PprDev &p) {
    free_dev(p.X[0]);
    free_dev(p.X[1]);
    free_dev(p.C);
    free_dev(p.part);
    free_dev(p.tasks);
    free_dev(p.big);
    free_dev(p.src);
    free_dev(p.rows);
    p = PprDev{};
}

// In-edge view for PageRank gathers: the CSR of a symmetric graph, else the CSC.
static bool ppr_view(const mg_graph *g, const int **off, const int **adj) {
    if (g->flags & MG_GRAPH_SYMMETRIC) {
        *off = g->row_off;
        *adj = g->col;
        return true;
    }
    if (g->col_off) {
        *off = g->col_off;
        *adj = g->row_idx;
        return true;
    }
    return false;
}

int64_t ppr_dev_bytes(const mg_graph *g) {
    const int64_t n = g->n;
    const int64_t tasks = g->m / PPR_T + n / PPR_RUN + 2 * (g->m / PPR_T) + 64;
    return 3 * 4LL * PPR_SLOTS * n + 4LL * PPR_SLOTS * (g->m / PPR_T + 64) + (int64_t)sizeof(ChunkRec) * tasks +
           12LL * (g->m / PPR_T + 64) + 4LL * PPR_SLOTS + 4LL * PPR_SLOTS * n;
}

int ppr_dev_alloc(PprDev &p, mg_graph *g) {
    ppr_dev_free(p);
    const int *off = nullptr, *adj = nullptr;
    if (!ppr_view(g, &off, &adj))
        return set_error(MG_ERR_UNSUPPORTED, "personalized PageRank needs MG_GRAPH_SYMMETRIC or MG_GRAPH_BUILD_CSC");
    const int n = g->n;
    std::vector<int> h_off((size_t)n + 1);
    MG_HIP(copy_on(g->stream, h_off.data(), off, 4 * ((size_t)n + 1), hipMemcpyDeviceToHost));
    std::vector<ChunkRec> tasks;
    std::vector<int> big;
    int npart = 0;
    for (int v = 0; v < n;) {
        const int d = h_off[v + 1] - h_off[v];
        if (d > PPR_T) {
            const int np = (d + PPR_T - 1) / PPR_T;
            big.push_back(v);
            big.push_back(npart);
            big.push_back(np);
            for (int k = 0; k < np; ++k) {
                const int b = h_off[v] + k * PPR_T;
                tasks.push_back(ChunkRec{v, b, std::min(b + PPR_T, h_off[v + 1]), -(npart + k + 1)});
            }
            npart += np;
            ++v;
            continue;
        }
        int u = v, e = 0;
        while (u < n && u - v < PPR_RUN && h_off[u + 1] - h_off[u] <= PPR_T && e + (h_off[u + 1] - h_off[u]) <= PPR_T) {
            e += h_off[u + 1] - h_off[u];
            ++u;
        }
        if (u == v) ++u;  // cannot happen (d <= PPR_T), kept for termination
        tasks.push_back(ChunkRec{v, 0, 0, u - v});
        v = u;
    }
    p.n = n;
    p.ntasks = (int)tasks.size();
    p.nbig = (int)(big.size() / 3);
    p.npart = npart;
    hipError_t e = hipSuccess;
    auto A = [&](void **q, size_t bytes) {
        if (e == hipSuccess) e = alloc_refused() ? hipErrorOutOfMemory : hipMalloc(q, bytes ? bytes : 16);
    };
    A((void **)&p.X[0], 4 * (size_t)PPR_SLOTS * n);
    A((void **)&p.X[1], 4 * (size_t)PPR_SLOTS * n);
    A((void **)&p.C, 4 * (size_t)PPR_SLOTS * n);
    A((void **)&p.part, 4 * (size_t)PPR_SLOTS * npart);
    A((void **)&p.tasks, sizeof(ChunkRec) * tasks.size());
    A((void **)&p.big, 4 * big.size());
    A((void **)&p.src, 4 * PPR_SLOTS);
    A((void **)&p.rows, 4 * (size_t)PPR_SLOTS * n);
    if (e == hipSuccess) e = copy_on(g->stream, p.tasks, tasks.data(), sizeof(ChunkRec) * tasks.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !big.empty()) e = copy_on(g->stream, p.big, big.data(), 4 * big.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        std::vector<int> none(PPR_SLOTS, -1);
        e = copy_on(g->stream, p.src, none.data(), 4 * PPR_SLOTS, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemsetAsync(p.X[0], 0, 4 * (size_t)PPR_SLOTS * n, g->stream);
    if (e == hipSuccess) e = hipMemsetAsync(p.X[1], 0, 4 * (size_t)PPR_SLOTS * n, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) {
        ppr_dev_free(p);
        return set_error(MG_ERR_OOM, "personalized PageRank state allocation failed: %s", hipGetErrorString(e));
    }
    p.bytes = ppr_dev_bytes(g);
    return MG_OK;
}

static PprArgs ppr_args(PprDev &p, mg_graph *g, float t) {
    PprArgs a{};
    ppr_view(g, &a.in_off, &a.in_adj);
    a.out_off = g->row_off;
    a.X = p.X[p.cur];
    a.Xn = p.X[p.cur ^ 1];
    a.C = p.C;
    a.part = p.part;
    a.tasks = p.tasks;
    a.big = p.big;
    a.src = p.src;
    a.n = p.n;
    a.ntasks = p.ntasks;
    a.nbig = p.nbig;
    a.t = t;
    return a;
}

int ppr_seed(PprDev &p, hipStream_t st, const SeedList &sl, const SlotList &retire) {
    if (sl.n > 0)
        hipLaunchKernelGGL(ppr_seed_cols, dim3(grid_blocks(p.n, BLK)), dim3(BLK), 0, st, p.X[p.cur], p.n, sl);
    hipLaunchKernelGGL(ppr_set_src, dim3(1), dim3(PPR_SLOTS), 0, st, p.src, sl, retire);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int ppr_iterate(PprDev &p, mg_graph *g, hipStream_t st, float t, int iters) {
    for (int i = 0; i < iters; ++i) {
        PprArgs a = ppr_args(p, g, t);
        const int G = grid_blocks(std::max<int64_t>((int64_t)p.ntasks * WAVE, (int64_t)p.n * 4), BLK);
        hipLaunchKernelGGL(ppr_contrib, dim3(G), dim3(BLK), 0, st, a);
        hipLaunchKernelGGL(ppr_gather, dim3(G), dim3(BLK), 0, st, a);
        if (p.nbig) hipLaunchKernelGGL(ppr_combine, dim3(grid_blocks((int64_t)p.nbig * WAVE, BLK)), dim3(BLK), 0, st, a);
        MG_HIP(hipGetLastError());
        p.cur ^= 1;
    }
    return MG_OK;
}

int ppr_extract_rows(PprDev &p, hipStream_t st, const SlotList &sl) {
    if (sl.n <= 0) return MG_OK;
    hipLaunchKernelGGL(ppr_extract, dim3(grid_blocks((int64_t)((p.n + 63) / 64) * BLK, BLK)), dim3(BLK), 0, st,
                       (const float *)p.X[p.cur], p.rows, p.n, sl);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int64_t ppr_dev_size(const PprDev *p) { return p->bytes; }

void ppr_release(mg_graph *g) {
    if (g->steps && g->steps->pdev) {
        ppr_dev_free(*g->steps->pdev);
        delete g->steps->pdev;
        g->steps->pdev = nullptr;
    }
}

int ppr_iters_of(const mg_options &o) { return o.ppr_iters > 0 ? o.ppr_iters : 20; }

}  // namespace mg

using namespace mg;

extern "C" int mg_ppr_batch(mg_graph *g, const int32_t *srcs, int32_t k, float *rank_out, const mg_options *opt,
                            mg_stats *stats) {
    CallTimer timer;
    mg_options o;
    int rc = begin_search(g, opt, &o);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g->mu);
    if (k < 1 || k > PPR_SLOTS || !srcs)
        return set_error(MG_ERR_INVALID_ARG, "mg_ppr_batch: k must be in [1, %d] with a source array", PPR_SLOTS);
    if ((rc = check_batch_sources("mg_ppr_batch", g, srcs, k))) return rc;
    if (!(o.pr_teleport >= 0.f && o.pr_teleport <= 1.f))
        return set_error(MG_ERR_INVALID_ARG, "mg_ppr_batch: pr_teleport must be in [0, 1]");
    PprDev *&p = g->steps->pdev;
    if (!p) {
        p = new (std::nothrow) PprDev();
        if (!p) return set_error(MG_ERR_OOM, "host allocation failed");
        if ((rc = ppr_dev_alloc(*p, g))) {
            delete p;
            p = nullptr;
            return rc;
        }
    }
    SeedList sl{};
    SlotList retire{}, all{};
    retire.n = PPR_SLOTS;
    for (int s = 0; s < PPR_SLOTS; ++s) retire.slot[s] = s;
    sl.n = all.n = k;
    for (int s = 0; s < k; ++s) {
        sl.slot[s] = all.slot[s] = s;
        sl.src[s] = srcs[s];
    }
    const int iters = ppr_iters_of(o);
    MG_HIP(hipEventRecord(g->ev0, g->stream));
    if ((rc = ppr_seed(*p, g->stream, sl, retire))) return rc;
    if ((rc = ppr_iterate(*p, g, g->stream, o.pr_teleport, iters))) return rc;
    if ((rc = ppr_extract_rows(*p, g->stream, all))) return rc;
    MG_HIP(hipEventRecord(g->ev1, g->stream));
    MG_HIP(hipEventSynchronize(g->ev1));
    float ms = 0;
    MG_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->iterations = iters;
        stats->edges_traversed = (int64_t)k * g->m * iters;
        stats->vertices_reached = (int64_t)k * g->n;
        stats->device_ms = ms;
    }
    if ((rc = deliver(g, rank_out, p->rows, 4 * (size_t)k * g->n, o.out_on_device))) return rc;
    if (stats) stats->host_ms = timer.ms();
    return MG_OK;
}
