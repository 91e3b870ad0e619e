// Batched multi-source BFS: up to 256 sources share every edge read.
//
// The reference runs one search per source, re-initialising between them
// (Algorithms/BFS/bfs.cu:402-417).  Here each vertex carries three 256-bit masks
// (one 32-byte sector each; bit s = source s): V (visited), F (frontier), N (next).
// One level is
//   * push (sparse levels): every active vertex v (F[v] != 0, listed in the active
//     bitmap A[L%2]) ORs F[v] & ~V[u] into N[u] for each out-neighbour u (atomicOr on
//     the words that gain bits) and marks u in A[(L+1)%2]; lists longer than LANE_T
//     become TASK-edge tasks walked by one wave;
//   * pull (dense levels, needs in-edges): every vertex u whose mask is not yet full
//     ORs the frontier masks of its in-neighbours, stopping once every source has
//     reached it, and stores the new bits in N[u] (no atomics);
//   * finalize: F[v] = 0 for the old frontier, then for every marked u: new = N[u],
//     V[u] |= new, F[u] = new, N[u] = 0, and depth[s][u] = L + 1 for each new bit s.
// Depths are unique per source, so row s equals a single-source search from
// srcs[s].  Level close, direction choice and termination run on the device in the
// next step's single-workgroup kernel; steps are replayed from a captured HIP graph.
#include "batch.h"

namespace mg {

struct BatchArgs {
    Ctl *ctl;
    const int *off, *adj;        // push view (CSR)
    const int *in_off, *in_adj;  // pull view (in-edges), nullptr when none
    uint4 *V, *F, *N;            // 2 uint4 per vertex
    uint32_t *act[2];            // active-vertex bitmaps
    int *out;                    // k x n depths
    int *hard;                   // pull: vertices handed to a wave
    ChunkRec *tasks;
    int64_t task_cap;
    LevelPart *part;
    M8 full;                     // bits of the k sources
    int n, nwords, k, G;
};

// push: v's frontier bits that u has not seen yet go to N[u]
__device__ __forceinline__ void push_edge(const BatchArgs &a, const M8 &fv, int u, uint32_t *act_next) {
    const M8 nw = andnot8(fv, ld8(a.V, u));
    if (!any8(nw)) return;
    uint32_t *nu = reinterpret_cast<uint32_t *>(a.N + 2 * u);
#pragma unroll
    for (int w = 0; w < 8; ++w) {
        const uint32_t b = word8(nw, w);
        if (b) atomicOr(&nu[w], b);
    }
    const uint32_t bit = 1u << (u & 31);
    if (!(act_next[u >> 5] & bit)) atomicOr(&act_next[u >> 5], bit);
}

__global__ void batch_init(BatchArgs a, int iter_limit, int can_pull, int alpha, int beta, long long m) {
    const int64_t kn = (int64_t)a.k * a.n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < kn; i += (int64_t)gridDim.x * blockDim.x)
        a.out[i] = -1;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 2 * (int64_t)a.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        a.V[i] = make_uint4(0, 0, 0, 0);
        a.F[i] = make_uint4(0, 0, 0, 0);
        a.N[i] = make_uint4(0, 0, 0, 0);
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.nwords; i += gridDim.x * blockDim.x) {
        a.act[0][i] = 0;
        a.act[1][i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Ctl *c = a.ctl;
        ctl_reset(c, a.n, iter_limit, a.k);
        c->can_bu = can_pull;
        c->m_unexplored = m;  // pull once the frontier's out-edges exceed m / alpha
        c->alpha = alpha;
        c->beta = beta;
    }
}

// One thread per source: visited/frontier bit, depth 0, and the source's out-degree
// toward the traversed-edge count.
__global__ void batch_seed(BatchArgs a, const int *srcs) {
    const int s = threadIdx.x;
    if (s >= a.k) return;
    const int v = srcs[s];
    uint32_t *vv = reinterpret_cast<uint32_t *>(a.V + 2 * v);
    uint32_t *fv = reinterpret_cast<uint32_t *>(a.F + 2 * v);
    atomicOr(&vv[s >> 5], 1u << (s & 31));
    atomicOr(&fv[s >> 5], 1u << (s & 31));
    atomicOr(&a.act[0][v >> 5], 1u << (v & 31));
    a.out[(int64_t)s * a.n + v] = 0;
    atomicAdd((unsigned long long *)&a.ctl->edges_traversed, (unsigned long long)(a.off[v + 1] - a.off[v]));
    atomicAdd((unsigned long long *)&a.ctl->reached, 1ULL);
}

// One workgroup: closes the pending level (partials of batch_fin_b).
__global__ __launch_bounds__(SMALL_BLK) void batch_close(BatchArgs a) {
    Ctl *c = a.ctl;
    if (!c->pending) return;
    long long nf = 0, mf = 0, te = 0, nr = 0;
    for (int i = threadIdx.x; i < a.G; i += SMALL_BLK) {
        const LevelPart p = a.part[i];
        nf += p.nf;
        mf += p.mf;
        te += p.ex;
        nr += p.pad;
    }
    nf = block_sum<SMALL_BLK>(nf);
    mf = block_sum<SMALL_BLK>(mf);
    te = block_sum<SMALL_BLK>(te);
    nr = block_sum<SMALL_BLK>(nr);
    if (threadIdx.x == 0) {
        const int L = c->level;
        c->edges_traversed += te;
        c->reached += nr;
        int dir = c->dir;
        if (dir == DIR_BU) c->bu_levels++;
        if (c->can_bu) {
            if (dir == DIR_TD && mf * c->alpha > c->m_unexplored) dir = DIR_BU;
            else if (dir == DIR_BU && mf * c->beta < c->m_unexplored) dir = DIR_TD;
        }
        c->dir = dir;
        c->pending = 0;
        c->nchunks = 0;
        c->nhard = 0;
        c->level = L + 1;
        c->nf_prev = (int)nf;
        if (nf == 0 || L + 1 >= c->iter_limit) c->done = 1;
    }
}

__global__ __launch_bounds__(BLK) void batch_a(BatchArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const int L = c->level;
    uint32_t *act_next = a.act[(L + 1) & 1];
    if (c->dir == DIR_TD) {
        __shared__ int s_w[BLK / WAVE];
        __shared__ int s_tb;
        for_each_bit(a.act[L & 1], a.nwords, [&](int v) {
            int beg = 0, deg = 0;
            if (v >= 0) {
                beg = a.off[v];
                deg = a.off[v + 1] - beg;
            }
            const int nt = deg > LANE_T ? (deg + TASK - 1) / TASK : 0;
            int tot;
            const int tex = block_exclusive_scan<BLK>(nt, s_w, &tot);
            if (tot > 0) {
                if (threadIdx.x == 0) s_tb = atomicAdd(&c->nchunks, tot);
                __syncthreads();
                for (int t = 0; t < nt; ++t) {
                    const int64_t slot = (int64_t)s_tb + tex + t;
                    if (slot < a.task_cap)
                        a.tasks[slot] = ChunkRec{v, beg + t * TASK, min(beg + (t + 1) * TASK, beg + deg), 0};
                    else
                        atomicOr(&c->changed, 1 << 30);
                }
                __syncthreads();
            }
            if (nt == 0 && deg > 0) {
                const M8 fv = ld8(a.F, v);
                for (int e = beg; e < beg + deg; ++e) push_edge(a, fv, a.adj[e], act_next);
            }
        });
    } else {
        __shared__ int s_w[BLK / WAVE];
        __shared__ int s_hb;
        const int lane = lane_id();
        for (int base = blockIdx.x * BLK; base < a.n; base += gridDim.x * BLK) {
            const int u = base + (int)threadIdx.x;
            bool found = false, hard = false;
            if (u < a.n) {
                const M8 vu = ld8(a.V, u);
                if (any8(andnot8(a.full, vu))) {
                    M8 acc{make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)};
                    const int beg = a.in_off[u], end = a.in_off[u + 1];
                    const int lim = min(end, beg + BU_INLINE);
                    int e = beg;
                    for (; e < lim; ++e) {
                        acc = or8(acc, ld8(a.F, a.in_adj[e]));
                        if (!any8(andnot8(a.full, or8(acc, vu)))) break;
                    }
                    const M8 nw = andnot8(acc, vu);
                    if (any8(nw)) {
                        st8(a.N, u, nw);
                        found = true;
                    }
                    hard = e == lim && lim < end && any8(andnot8(a.full, or8(acc, vu)));
                }
            }
            const uint64_t fm = wave_ballot(found);
            const uint32_t mine = (lane < 32) ? (uint32_t)fm : (uint32_t)(fm >> 32);
            if ((lane & 31) == 0 && mine) act_next[u >> 5] = mine;  // each wave owns its 64 vertices' words
            int htot;
            const int hex = block_exclusive_scan<BLK>(hard ? 1 : 0, s_w, &htot);
            if (htot > 0) {
                if (threadIdx.x == 0) s_hb = atomicAdd(&c->nhard, htot);
                __syncthreads();
                if (hard) a.hard[s_hb + hex] = u;
                __syncthreads();
            }
        }
    }
}

__global__ __launch_bounds__(BLK) void batch_b(BatchArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const int L = c->level;
    uint32_t *act_next = a.act[(L + 1) & 1];
    const int lane = lane_id();
    const int wave = (blockIdx.x * BLK + (int)threadIdx.x) / WAVE, nwaves = gridDim.x * BLK / WAVE;
    if (c->dir == DIR_TD) {
        const int nt = (int)min((int64_t)c->nchunks, a.task_cap);
        for (int t = wave; t < nt; t += nwaves) {
            const ChunkRec r = a.tasks[t];
            const M8 fv = ld8(a.F, r.v);
            for (int e = r.beg + lane; e < r.end; e += WAVE) push_edge(a, fv, a.adj[e], act_next);
        }
    } else {
        const int nh = c->nhard;
        for (int h = wave; h < nh; h += nwaves) {
            const int u = a.hard[h];
            const M8 vu = ld8(a.V, u);
            const M8 prev = ld8(a.N, u);  // bits found by the lane scan
            M8 acc{make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)};
            const int end = a.in_off[u + 1];
            for (int e = a.in_off[u] + BU_INLINE + lane; e < end; e += WAVE) acc = or8(acc, ld8(a.F, a.in_adj[e]));
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) acc = or8(acc, shfl_xor8(acc, o));
            if (lane == 0) {
                const M8 nw = andnot8(or8(acc, prev), vu);
                if (any8(nw)) {
                    st8(a.N, u, nw);
                    const uint32_t bit = 1u << (u & 31);
                    atomicOr(&act_next[u >> 5], bit);
                }
            }
        }
    }
}

// Finalize, part 1: the old frontier's masks are cleared.
__global__ __launch_bounds__(BLK) void batch_fin_a(BatchArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const int L = c->level;
    for_each_bit(a.act[L & 1], a.nwords, [&](int v) {
        if (v >= 0) st8(a.F, v, M8{make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)});
    });
}

// Finalize, part 2: new bits become visited and the next frontier; depths written.
__global__ __launch_bounds__(BLK) void batch_fin_b(BatchArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const int L = c->level;
    long long nf = 0, mf = 0, te = 0, nr = 0;
    // the old active bitmap becomes the next level's (empty) target
    uint32_t *old = a.act[L & 1];
    for (int i = blockIdx.x * BLK + (int)threadIdx.x; i < a.nwords; i += gridDim.x * BLK) old[i] = 0;
    for_each_bit(a.act[(L + 1) & 1], a.nwords, [&](int u) {
        if (u < 0) return;
        const M8 nw = ld8(a.N, u);
        st8(a.N, u, M8{make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)});
        st8(a.V, u, or8(ld8(a.V, u), nw));
        st8(a.F, u, nw);
        const int deg = a.off[u + 1] - a.off[u];
        const int cnt = popc8(nw);
        ++nf;
        nr += cnt;
        mf += deg;
        te += (long long)cnt * deg;
        for (int w = 0; w < 8; ++w) {
            uint32_t b = word8(nw, w);
            while (b) {
                const int s = w * 32 + __ffs(b) - 1;
                a.out[(int64_t)s * a.n + u] = L + 1;
                b &= b - 1;
            }
        }
    });
    nf = block_sum<BLK>(nf);
    mf = block_sum<BLK>(mf);
    te = block_sum<BLK>(te);
    nr = block_sum<BLK>(nr);
    if (threadIdx.x == 0) {
        LevelPart p;
        p.nf = nf;
        p.mf = mf;
        p.ex = te;
        p.pad = nr;  // (source, vertex) pairs reached
        a.part[blockIdx.x] = p;
        if (blockIdx.x == 0) c->pending = 1;
    }
}

void batch_release(mg_graph *g) {
    BatchState &bs = g->steps->batch;
    free_dev(bs.V);
    free_dev(bs.F);
    free_dev(bs.N);
    free_dev(bs.out);
    free_dev(bs.srcs);
    free_dev(bs.D);
    free_dev(bs.tasks);
    free_dev(bs.FAR);
    free_dev(bs.fb);
    bs = BatchState{};
}

// Allocates the batch state of graph g once, at the maximum batch size, so that
// captured step graphs keep valid pointers for the graph's lifetime.
int ensure_batch(mg_graph *g, bool sssp) {
    BatchState &bs = g->steps->batch;
    const size_t n = (size_t)g->n;
    hipError_t e = hipSuccess;
    if (!bs.V) {
        e = hipMalloc(&bs.V, 32 * n + 32);
        if (e == hipSuccess) e = hipMalloc(&bs.F, 32 * n + 32);
        if (e == hipSuccess) e = hipMalloc(&bs.N, 32 * n + 32);
        if (e == hipSuccess) e = hipMalloc((void **)&bs.out, 4 * (size_t)MG_BATCH_MAX * n + 4);
        if (e == hipSuccess) e = hipMalloc((void **)&bs.srcs, 4 * MG_BATCH_MAX);
        if (e != hipSuccess) {
            batch_release(g);
            return set_error(MG_ERR_OOM, "batch state allocation failed: %s", hipGetErrorString(e));
        }
        bs.n = g->n;
        bs.out_cap = (size_t)MG_BATCH_MAX * n;
    }
    if (sssp && !bs.D) {
        bs.task_cap = g->m / SB_TASK + (int64_t)n + 64;
        e = hipMalloc((void **)&bs.D, 4 * (size_t)MG_BATCH_MAX * n + 64);
        if (e == hipSuccess) e = hipMalloc((void **)&bs.tasks, sizeof(ChunkRec) * (size_t)bs.task_cap);
        if (e == hipSuccess) e = hipMalloc(&bs.FAR, 32 * n + 32);
        if (e == hipSuccess) e = hipMalloc((void **)&bs.fb, 4 * (size_t)SB_RING * g->ws.nwords);
        if (e != hipSuccess) {
            free_dev(bs.D);
            free_dev(bs.tasks);
            free_dev(bs.FAR);
            free_dev(bs.fb);
            bs.D = nullptr;
            bs.tasks = nullptr;
            bs.FAR = nullptr;
            bs.fb = nullptr;
            return set_error(MG_ERR_OOM, "batched SSSP state allocation failed: %s", hipGetErrorString(e));
        }
    }
    return MG_OK;
}

int check_batch_sources(const char *fn, const mg_graph *g, const int32_t *srcs, int32_t k) {
    if (k < 1 || k > MG_BATCH_MAX || !srcs)
        return set_error(MG_ERR_INVALID_ARG, "%s: k must be in [1, %d] with a source array", fn, MG_BATCH_MAX);
    for (int i = 0; i < k; ++i)
        if (srcs[i] < 0 || srcs[i] >= g->n)
            return set_error(MG_ERR_INVALID_ARG, "%s: source %d = %d outside [0,%d)", fn, i, srcs[i], g->n);
    return MG_OK;
}

int batch_alpha(const mg_options &o) { return o.batch_alpha > 0 ? o.batch_alpha : 4; }
int batch_beta(const mg_options &o) { return o.batch_beta > 0 ? o.batch_beta : 16; }

}  // namespace mg

using namespace mg;

extern "C" int mg_bfs_batch(mg_graph *g, const int32_t *srcs, int32_t k, int32_t *depth_out, const mg_options *opt,
                            mg_stats *stats) {
    CallTimer timer;
    mg_options o;
    int rc = begin_search(g, opt, &o);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g->mu);
    if ((rc = check_batch_sources("mg_bfs_batch", g, srcs, k))) return rc;
    if (o.directed == 0 && !(g->flags & MG_GRAPH_SYMMETRIC))
        return set_error(MG_ERR_UNSUPPORTED, "mg_bfs_batch: directed=0 needs a symmetric graph");
    if ((rc = ensure_batch(g, false))) return rc;
    Workspace &ws = g->ws;
    const int n = g->n;
    BatchState &bs = g->steps->batch;
    BatchArgs a{};
    a.ctl = ws.ctl;
    a.off = g->row_off;
    a.adj = g->col;
    if (g->flags & MG_GRAPH_SYMMETRIC) {
        a.in_off = g->row_off;
        a.in_adj = g->col;
    } else if (g->col_off) {
        a.in_off = g->col_off;
        a.in_adj = g->row_idx;
    }
    a.V = reinterpret_cast<uint4 *>(bs.V);
    a.F = reinterpret_cast<uint4 *>(bs.F);
    a.N = reinterpret_cast<uint4 *>(bs.N);
    a.act[0] = ws.bm[0];
    a.act[1] = ws.bm[1];
    a.out = bs.out;
    a.hard = ws.hard;
    a.tasks = ws.chunks;
    a.task_cap = ws.chunk_cap;
    a.part = ws.part;
    uint32_t fw[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int s = 0; s < k; ++s) fw[s >> 5] |= 1u << (s & 31);
    a.full = M8{make_uint4(fw[0], fw[1], fw[2], fw[3]), make_uint4(fw[4], fw[5], fw[6], fw[7])};
    a.n = n;
    a.nwords = ws.nwords;
    a.k = k;
    a.G = grid_blocks(std::max<int64_t>(n / 4, g->m / 64), BLK);
    const int can_pull = a.in_off != nullptr;

    MG_HIP(hipMemcpyAsync(bs.srcs, srcs, 4 * (size_t)k, hipMemcpyHostToDevice, g->stream));
    MG_HIP(hipEventRecord(g->ev0, g->stream));
    hipLaunchKernelGGL(batch_init, dim3(grid_blocks((int64_t)k * n, BLK)), dim3(BLK), 0, g->stream, a, o.iter_num,
                       can_pull, batch_alpha(o), batch_beta(o), (long long)g->m);
    hipLaunchKernelGGL(batch_seed, dim3(1), dim3(MG_BATCH_MAX), 0, g->stream, a, (const int *)bs.srcs);
    MG_HIP(hipGetLastError());
    auto step = [&]() -> int {
        hipLaunchKernelGGL(batch_close, dim3(1), dim3(SMALL_BLK), 0, g->stream, a);
        hipLaunchKernelGGL(batch_a, dim3(a.G), dim3(BLK), 0, g->stream, a);
        hipLaunchKernelGGL(batch_b, dim3(a.G), dim3(BLK), 0, g->stream, a);
        hipLaunchKernelGGL(batch_fin_a, dim3(a.G), dim3(BLK), 0, g->stream, a);
        hipLaunchKernelGGL(batch_fin_b, dim3(a.G), dim3(BLK), 0, g->stream, a);
        MG_HIP(hipGetLastError());
        return MG_OK;
    };
    // the captured steps bake in k and the full mask
    const uint64_t key = mix64((uint64_t)k * 2 + (uint64_t)o.directed);
    const int64_t max_steps = std::min<int64_t>((int64_t)o.iter_num, (int64_t)n) + 3;
    if ((rc = drive(g, step, max_steps, KIND_BATCH, key))) return rc;
    rc = finish_search(g, stats, nullptr, 0, depth_out, bs.out, 4 * (size_t)k * n, o.out_on_device, timer);
    if (rc) return rc;
    if (stats) {
        stats->edges_traversed = ws.h_ctl->edges_traversed;
        stats->vertices_reached = ws.h_ctl->reached;
    }
    return MG_OK;
}
