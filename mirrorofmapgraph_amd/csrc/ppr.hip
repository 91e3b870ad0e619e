// Batched personalized PageRank: up to 64 sources (slots) iterate together.
//
// Not in the reference, whose PageRank (Algorithms/PageRank/PageRank.h:233-283) has
// one global teleport term; this is its personalized form, with the reference's
// conventions: contribution r[v] / outdeg(v) gathered over in-edges, rank-sink
// (dangling) mass dropped, teleport t applied as written there (t + (1 - t) * sum):
//   x_0 = e_s,   x_{i+1}[u] = (1 - t) * sum_{v -> u} x_i[v] / outdeg(v)  (+ t when u = s).
// Ranks are vertex-major, X[v * W + slot], W = 8, 16, 32 or 64 floats per vertex: the
// smallest width above every slot in use (slots are handed out lowest first; the rows
// are repacked when it changes), so a wave gathers one in-edge's row for 64 / W edges
// at a time and a row is one 32-byte sector while at most 8 slots are in use.
// Vertices are covered by static tasks: runs of consecutive vertices with at most
// PPR_T in-edges in total (one wave walks them), and PPR_T-edge pieces of longer
// in-lists whose partial sums are added in piece order.  Every list is summed in one
// order fixed by the list alone, whatever the width (see gather_sum), so a slot's
// arithmetic is fixed by the graph: a query gives the same bits alone or beside any
// other queries.
#include "batch.h"

#include <vector>

namespace mg {

constexpr int PPR_T = 1024;    // in-edges per task
constexpr int PPR_RUN = 64;    // vertices per run task

struct PprArgs {
    const int *in_off, *in_adj;  // in-edges (CSR of a symmetric graph, or the CSC)
    const int *out_off;          // out-degrees
    const float *X;              // current ranks, n x W
    float *Xn;                   // next ranks
    float *C;                    // contributions X / outdeg
    float *part;                 // piece partial sums, npart x W
    const ChunkRec *tasks;       // runs: {v0, -, -, count}; pieces: {v, beg, end, -(part + 1)}
    const int *big;              // per long list: vertex, first part, parts (3 ints)
    const int *src;              // per slot: source or -1
    int n, ntasks, nbig;
    float t;                     // teleport
};

// Row width: ranks, contributions and partial sums are stored n x W, W the smallest
// of 8, 16, 32, 64 above every slot in use (slots are handed out lowest first), so a
// gathered row is W floats: 32 bytes (one sector) while at most 8 slots are in use,
// and the contribution array of a 4M-vertex graph is 128 MB instead of 1 GB.
__host__ __device__ constexpr int ppr_groups(int W) { return WAVE / W; }

// C[v][c] = X[v][c] / outdeg(v) (0 for dangling v), over all n x W entries.
template <int W>
__global__ __launch_bounds__(BLK) void ppr_contrib(PprArgs a) {
    const int64_t total = (int64_t)a.n * W;
    for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < total; i += (int64_t)gridDim.x * BLK) {
        const int v = (int)(i / W);
        const int d = a.out_off[v + 1] - a.out_off[v];
        a.C[i] = d > 0 ? a.X[i] / (float)d : 0.f;
    }
}

// Sum of the lane's column over in-edges [beg, end), in one order for every width:
// the edge at offset k of the list goes to partial sum k mod 8 (each partial sum adds
// its edges in list order), and the eight partial sums are combined as
// ((s0 + s1) + (s2 + s3)) + ((s4 + s5) + (s6 + s7)).  A wave holds G = 64 / W edge
// groups of W lanes (lane = g * W + column): group g owns partial sums g + G j
// (j < 8 / G), so a wave instruction gathers G rows; the combination steps over
// partial sums on different groups are lane butterflies (xor W, 2W, 4W), the others
// adds inside the lane, and addition is commutative, so every width and every lane of
// a pair computes the same bits.  Eight gathers per lane are in flight per step
// (G rounds of eight edges); the last step loads from clamped slots and adds only
// the edges in range.
template <int W>
__device__ __forceinline__ float gather_sum(const PprArgs &a, int beg, int end, int g, int col) {
    constexpr int G = ppr_groups(W), J = 8 / G;
    float acc[J];
#pragma unroll
    for (int j = 0; j < J; ++j) acc[j] = 0.f;
    for (int e0 = beg; e0 < end; e0 += 8 * G) {
        // round r of the step covers edges e0 + 8 r .. e0 + 8 r + 7; this lane's edge
        // of partial sum g + G j in round r is e0 + 8 r + g + G j
        int u[G][J];
        bool ok[G][J];
#pragma unroll
        for (int r = 0; r < G; ++r)
#pragma unroll
            for (int j = 0; j < J; ++j) {
                const int e = e0 + 8 * r + g + G * j;
                ok[r][j] = e < end;
                u[r][j] = a.in_adj[ok[r][j] ? e : beg];
            }
        float c[G][J];
#pragma unroll
        for (int r = 0; r < G; ++r)
#pragma unroll
            for (int j = 0; j < J; ++j) c[r][j] = a.C[(int64_t)u[r][j] * W + col];
#pragma unroll
        for (int r = 0; r < G; ++r)
#pragma unroll
            for (int j = 0; j < J; ++j)
                if (ok[r][j]) acc[j] += c[r][j];
    }
    // combine: level b pairs partial sums whose index differs in bit b
    float s[J];
#pragma unroll
    for (int j = 0; j < J; ++j) s[j] = acc[j];
    int nj = J;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        if ((1 << b) < G) {  // partner on group g ^ (1 << b)
#pragma unroll
            for (int j = 0; j < J; ++j) s[j] += __shfl_xor(s[j], W << b);
        } else {  // partner in this lane: j ^ (1 << (b - log2 G))
            const int h = nj / 2;
#pragma unroll
            for (int j = 0; j < J / 2; ++j)
                if (j < h) s[j] = s[2 * j] + s[2 * j + 1];
            nj = h;
        }
    }
    return s[0];
}

// The same sum for a list of at most 8 edges, by one group alone: edge k is partial
// sum k (the others are 0), combined in the same tree inside the lane.  Slots past
// the end load from beg, which is in range only for a non-empty list (an empty list
// at the end of the array has beg = m): an empty list loads nothing and sums to 0,
// as the tree of eight zero partial sums does.
template <int W>
__device__ __forceinline__ float short_sum(const PprArgs &a, int beg, int end, int col) {
    if (end <= beg) return 0.f;
    int u[8];
    float c[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) u[k] = a.in_adj[beg + k < end ? beg + k : beg];
#pragma unroll
    for (int k = 0; k < 8; ++k) c[k] = a.C[(int64_t)u[k] * W + col];
#pragma unroll
    for (int k = 0; k < 8; ++k) c[k] = beg + k < end ? 0.f + c[k] : 0.f;
    return ((c[0] + c[1]) + (c[2] + c[3])) + ((c[4] + c[5]) + (c[6] + c[7]));
}

__device__ __forceinline__ float ppr_apply(float s, float t, bool is_src) {
    const float v = (1.0f - t) * s;
    return is_src ? t + v : v;
}

template <int W>
__global__ __launch_bounds__(BLK) void ppr_gather(PprArgs a) {
    constexpr int G = ppr_groups(W);
    const int lane = lane_id(), g = lane / W, col = lane % W;
    const int wave = (blockIdx.x * BLK + (int)threadIdx.x) / WAVE, nwaves = gridDim.x * BLK / WAVE;
    const int src = a.src[col];
    const bool on = src >= 0;
    if (!__any(on)) return;
    for (int t = wave; t < a.ntasks; t += nwaves) {
        const ChunkRec r = a.tasks[t];
        if (r.pad > 0) {
            // G vertices of the run at a time: lists of at most 8 edges are summed by
            // their group alone, longer ones by the whole wave, one after another
            // (idle columns unused)
            for (int u0 = r.v; u0 < r.v + r.pad; u0 += G) {
                const int u = u0 + g;
                const bool has = u < r.v + r.pad;
                const int b = has ? a.in_off[u] : 0, e = has ? a.in_off[u + 1] : 0;
                const bool shortl = has && e - b <= 8;
                if (__any(shortl) && shortl) {
                    const float s = short_sum<W>(a, b, e, col);
                    if (on) a.Xn[(int64_t)u * W + col] = ppr_apply(s, a.t, u == src);
                }
                uint64_t longs = wave_ballot(has && !shortl && col == 0);  // bit g * W per long list
                while (longs) {
                    const int gl = (__ffsll((unsigned long long)longs) - 1) / W;
                    longs &= longs - 1;
                    const int ul = u0 + gl;
                    const float s = gather_sum<W>(a, a.in_off[ul], a.in_off[ul + 1], g, col);
                    if (on && g == 0) a.Xn[(int64_t)ul * W + col] = ppr_apply(s, a.t, ul == src);
                }
            }
        } else {
            const float s = gather_sum<W>(a, r.beg, r.end, g, col);
            if (on && g == 0) a.part[(int64_t)(-r.pad - 1) * W + col] = s;
        }
    }
}

// Long in-lists: the pieces' partial sums added in piece order (G lists per wave).
template <int W>
__global__ __launch_bounds__(BLK) void ppr_combine(PprArgs a) {
    constexpr int G = ppr_groups(W);
    const int lane = lane_id(), g = lane / W, col = lane % W;
    const int wave = (blockIdx.x * BLK + (int)threadIdx.x) / WAVE, nwaves = gridDim.x * BLK / WAVE;
    const int src = a.src[col];
    const bool on = src >= 0;
    for (int b0 = wave * G; b0 < a.nbig; b0 += nwaves * G) {
        const int b = b0 + g;
        if (b >= a.nbig) continue;
        const int u = a.big[3 * b], p0 = a.big[3 * b + 1], np = a.big[3 * b + 2];
        float s = 0.f;
        for (int p = p0; p < p0 + np; ++p) s += on ? a.part[(int64_t)p * W + col] : 0.f;
        if (on) a.Xn[(int64_t)u * W + col] = ppr_apply(s, a.t, u == src);
    }
}

// Rows of width w1 -> rows of width w2 (columns beyond w1 start at 0).
__global__ void ppr_repack(const float *X, int w1, float *Y, int w2, int n) {
    const int64_t total = (int64_t)n * w2;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = i / w2;
        const int c = (int)(i % w2);
        Y[i] = c < w1 ? X[v * w1 + c] : 0.f;
    }
}

// Column reset of the listed slots: x_0 = e_src.
__global__ void ppr_seed_cols(float *X, int n, int W, SeedList sl) {
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, nt = (int64_t)gridDim.x * blockDim.x;
    for (int64_t u = tid; u < n; u += nt)
        for (int j = 0; j < sl.n; ++j) X[u * W + sl.slot[j]] = u == sl.src[j] ? 1.f : 0.f;
}

__global__ void ppr_set_src(int *src, SeedList sl, SlotList off) {
    const int i = threadIdx.x;
    if (i < off.n) src[off.slot[i]] = -1;
    __syncthreads();
    if (i < sl.n) src[sl.slot[i]] = sl.src[i];
}

// Rank columns of the listed slots -> their rows (64 vertices x W columns through LDS).
__global__ __launch_bounds__(BLK) void ppr_extract(const float *X, int W, float *rows, int n, SlotList sl) {
    __shared__ float tile[64][65];
    const int tv = (n + 63) / 64;
    for (int t = blockIdx.x; t < tv; t += gridDim.x) {
        const int v0 = t * 64;
        for (int i = threadIdx.x; i < 64 * W; i += BLK) {
            const int r = i / W, cc = i % W;
            tile[r][cc] = (v0 + r < n) ? X[(int64_t)(v0 + r) * W + cc] : 0.f;
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

int ppr_dev_alloc(PprDev &p, mg_graph *g, hipStream_t st) {
    ppr_dev_free(p);
    const int *off = nullptr, *adj = nullptr;
    if (!ppr_view(g, &off, &adj))
        return set_error(MG_ERR_UNSUPPORTED, "personalized PageRank needs MG_GRAPH_SYMMETRIC or MG_GRAPH_BUILD_CSC");
    const int n = g->n;
    std::vector<int> h_off((size_t)n + 1);
    MG_HIP(copy_on(st, h_off.data(), off, 4 * ((size_t)n + 1), hipMemcpyDeviceToHost));
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
    if (e == hipSuccess) e = copy_on(st, p.tasks, tasks.data(), sizeof(ChunkRec) * tasks.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !big.empty()) e = copy_on(st, p.big, big.data(), 4 * big.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        std::vector<int> none(PPR_SLOTS, -1);
        e = copy_on(st, p.src, none.data(), 4 * PPR_SLOTS, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemsetAsync(p.X[0], 0, 4 * (size_t)PPR_SLOTS * n, st);
    if (e == hipSuccess) e = hipMemsetAsync(p.X[1], 0, 4 * (size_t)PPR_SLOTS * n, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        ppr_dev_free(p);
        return set_error(MG_ERR_OOM, "personalized PageRank state allocation failed: %s", hipGetErrorString(e));
    }
    p.bytes = ppr_dev_bytes(g);
    return MG_OK;
}

// The row width for the slots in `live`.
static int ppr_width(uint64_t live) {
    const int need = live ? 64 - __builtin_clzll(live) : 1;
    return need <= 8 ? 8 : need <= 16 ? 16 : need <= 32 ? 32 : 64;
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
    uint64_t live = p.live;
    for (int i = 0; i < retire.n; ++i) live &= ~(1ull << retire.slot[i]);
    for (int i = 0; i < sl.n; ++i) live |= 1ull << sl.slot[i];
    const int w = ppr_width(live);
    if (w != p.width) {  // rows of the new width into the other buffer
        hipLaunchKernelGGL(ppr_repack, dim3(grid_blocks_cap((int64_t)p.n * w, BLK, 4 * MAX_GRID)), dim3(BLK), 0, st,
                           (const float *)p.X[p.cur], p.width, p.X[p.cur ^ 1], w, p.n);
        p.cur ^= 1;
        p.width = w;
    }
    p.live = live;
    if (sl.n > 0)
        hipLaunchKernelGGL(ppr_seed_cols, dim3(grid_blocks(p.n, BLK)), dim3(BLK), 0, st, p.X[p.cur], p.n, p.width, sl);
    hipLaunchKernelGGL(ppr_set_src, dim3(1), dim3(PPR_SLOTS), 0, st, p.src, sl, retire);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

template <int W>
static int ppr_iterate_w(PprDev &p, mg_graph *g, hipStream_t st, float t, int iters) {
    for (int i = 0; i < iters; ++i) {
        PprArgs a = ppr_args(p, g, t);
        const int G = grid_blocks(std::max<int64_t>((int64_t)p.ntasks * WAVE, (int64_t)p.n * 4), BLK);
        hipLaunchKernelGGL(ppr_contrib<W>, dim3(grid_blocks_cap((int64_t)p.n * W, BLK, 4 * MAX_GRID)), dim3(BLK), 0,
                           st, a);
        hipLaunchKernelGGL(ppr_gather<W>, dim3(G), dim3(BLK), 0, st, a);
        if (p.nbig)
            hipLaunchKernelGGL(ppr_combine<W>, dim3(grid_blocks((int64_t)p.nbig * WAVE / ppr_groups(W), BLK) + 1),
                               dim3(BLK), 0, st, a);
        MG_HIP(hipGetLastError());
        p.cur ^= 1;
    }
    return MG_OK;
}

int ppr_iterate(PprDev &p, mg_graph *g, hipStream_t st, float t, int iters) {
    switch (p.width) {
    case 8: return ppr_iterate_w<8>(p, g, st, t, iters);
    case 16: return ppr_iterate_w<16>(p, g, st, t, iters);
    case 32: return ppr_iterate_w<32>(p, g, st, t, iters);
    default: return ppr_iterate_w<64>(p, g, st, t, iters);
    }
}

int ppr_extract_rows(PprDev &p, hipStream_t st, const SlotList &sl) {
    if (sl.n <= 0) return MG_OK;
    hipLaunchKernelGGL(ppr_extract, dim3(grid_blocks((int64_t)((p.n + 63) / 64) * BLK, BLK)), dim3(BLK), 0, st,
                       (const float *)p.X[p.cur], p.width, p.rows, p.n, sl);
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
    // the degree-ordered twin when the graph keeps one (sources renamed, rows mapped
    // back); its in-lists are in twin order, so the sums are the twin's, as for serving
    int32_t tsrc[PPR_SLOTS];
    mg_graph *sg = batch_graph(g, o, srcs, k, tsrc);
    if (sg != g) srcs = tsrc;
    PprDev *&p = sg->steps->pdev;
    if (!p) {
        p = new (std::nothrow) PprDev();
        if (!p) return set_error(MG_ERR_OOM, "host allocation failed");
        if ((rc = ppr_dev_alloc(*p, sg, g->stream))) {
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
    if ((rc = ppr_iterate(*p, sg, g->stream, o.pr_teleport, iters))) return rc;
    if ((rc = ppr_extract_rows(*p, g->stream, all))) return rc;
    const float *result = p->rows;
    if (sg != g) {  // rows 0..k-1 back to the caller's ids, into the spare rank buffer
        float *spare = p->X[p->cur ^ 1];
        if ((rc = launch_unpermute_rows(sg->perm, (const int *)p->rows, g->n, k, (int *)spare, g->stream))) return rc;
        result = spare;
    }
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
    if ((rc = deliver(g, rank_out, result, 4 * (size_t)k * g->n, o.out_on_device))) return rc;
    if (stats) stats->host_ms = timer.ms();
    return MG_OK;
}
