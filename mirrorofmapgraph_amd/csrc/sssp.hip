// Single-source shortest paths: the engine's specialised path for the sssp program
// (Algorithms/SSSP/sssp.h).
//
// The reference program is push-style frontier Bellman-Ford: expand emits
// dist[v] + w for out-edges that improve on dist[nbr] (sssp.h:342-357), contract
// atomicMin's it into gather_tmp (sssp.h:384-401), apply/post_apply commit the
// minimum and mark changed vertices (sssp.h:242-280), and changed vertices form the
// next frontier.  Here one fused pass per level relaxes an edge straight into dist
// with atomicMin and enqueues the target once (test-and-set on the next-frontier
// bitmap) when it improved; the level loop, frontier rotation and termination stay
// on the device.  Integer shortest distances are unique, so the result equals the
// reference's: dist[src] = 0, unreached = INT_MAX (sssp.h:42-43, 95-100).
#include <climits>

#include "frontier.h"

namespace mg {

struct SsspArgs {
    FrontierArgs fa;
    int *dist;
    const int *w;
};

struct SsspPolicy {
    const SsspArgs &a;
    long long mf = 0;   // out-edges of the vertices enqueued (small-frontier test)
    long long ex = 0;   // edges relaxed
    __device__ explicit SsspPolicy(const SsspArgs &args) : a(args) {}
    __device__ int load(int v) const { return a.dist[v]; }
    __device__ void visit(int view, int, int d, int slot, int u, bool valid, int L) {
        bool improved = false;
        if (valid) {
            ++ex;
            const long long nd = (long long)d + a.w[edge_id(a.fa, view, slot)];
            if (nd < (long long)a.dist[u]) improved = (int)nd < atomicMin(&a.dist[u], (int)nd);
        }
        if (enqueue_next(a.fa, L, u, improved)) mf += view_degree(a.fa, u);
    }
};

__device__ void sssp_level_end(const SsspArgs &a) {
    Ctl *c = a.fa.ctl;
    c->mf_cur = ld_dev(&c->mf_next);
    rotate_level(c, c->level);
}

__global__ void sssp_seed(SsspArgs a, int src, int iter_limit) {
    Ctl *c = a.fa.ctl;
    a.dist[src] = 0;
    a.fa.bm[0][src >> 5] = 1u << (src & 31);
    a.fa.q[0][0] = src;
    ctl_reset(c, a.fa.n, iter_limit, 1);
    c->mf_cur = view_degree(a.fa, src);
}

__device__ __forceinline__ void flush_counts(Ctl *c, long long mf, long long ex) {
    mf = wave_sum_ll(mf);
    ex = wave_sum_ll(ex);
    if (lane_id() == 0) {
        if (mf) atomicAdd((unsigned long long *)&c->mf_next, (unsigned long long)mf);
        if (ex) atomicAdd((unsigned long long *)&c->edges_examined, (unsigned long long)ex);
    }
}

__global__ __launch_bounds__(SMALL_BLK) void sssp_small(SsspArgs a) {
    Ctl *c = a.fa.ctl;
    __shared__ int s_go, s_L;
    for (;;) {
        if (threadIdx.x == 0) {
            const int L = c->level;
            s_go = !c->done && c->qsize[L % 3] <= SMALL_Q && c->mf_cur <= SMALL_E;
            s_L = L;
        }
        __syncthreads();
        if (!s_go) break;
        const int L = s_L;
        clear_prev<SMALL_BLK>(a.fa, L, 0, 1);
        __syncthreads();
        SsspPolicy pol(a);
        for (int view = 0; view < a.fa.nviews; ++view) td_tiles<SMALL_BLK, false>(a.fa, pol, view, L, 0, 1);
        flush_counts(c, pol.mf, pol.ex);
        __syncthreads();
        if (threadIdx.x == 0) {
            c->small_levels++;
            sssp_level_end(a);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(BLK) void sssp_level_a(SsspArgs a, int view) {
    Ctl *c = a.fa.ctl;
    if (c->done) return;
    const int L = c->level;
    if (view == 0) clear_prev<BLK>(a.fa, L, blockIdx.x, gridDim.x);
    SsspPolicy pol(a);
    td_tiles<BLK, true>(a.fa, pol, view, L, blockIdx.x, gridDim.x);
    flush_counts(c, pol.mf, pol.ex);
}

__global__ __launch_bounds__(BLK) void sssp_level_b(SsspArgs a) {
    Ctl *c = a.fa.ctl;
    if (c->done) return;
    SsspPolicy pol(a);
    td_chunks<BLK>(a.fa, pol, c->level);
    flush_counts(c, pol.mf, pol.ex);
    if (last_block(c)) {
        sssp_level_end(a);
        __threadfence();
    }
}

}  // namespace mg

using namespace mg;

extern "C" int mg_sssp(mg_graph *g, int32_t src, int32_t *dist_out, const mg_options *opt, mg_stats *stats) {
    CallTimer timer;
    mg_options o;
    int rc = begin_search(g, opt, &o);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g->mu);
    if (!g->w) return set_error(MG_ERR_UNSUPPORTED, "mg_sssp: graph has no edge weights (MG_GRAPH_WEIGHTED)");
    if (g->min_weight < 0)
        return set_error(MG_ERR_UNSUPPORTED, "mg_sssp: negative edge weight %d (weights must be >= 0)",
                         g->min_weight);
    if (src < 0 || src >= g->n)
        return set_error(MG_ERR_INVALID_ARG, "mg_sssp: source %d outside [0,%d)", src, g->n);
    Workspace &ws = g->ws;
    SsspArgs a{};
    a.fa = make_frontier_args(g, o.directed);
    a.dist = ws.vals;
    a.w = g->w;
    const int G = grid_blocks(std::max<int64_t>(g->n / 4, g->m / 64), BLK);

    MG_HIP(hipEventRecord(g->ev0, g->stream));
    if ((rc = launch_fill(g, ws.vals, INT_MAX, g->n))) return rc;
    for (int i = 0; i < 3; ++i)
        if ((rc = launch_fill(g, ws.bm[i], 0u, ws.nwords))) return rc;
    hipLaunchKernelGGL(sssp_seed, dim3(1), dim3(1), 0, g->stream, a, src, o.iter_num);
    MG_HIP(hipGetLastError());
    auto step = [&]() -> int {
        hipLaunchKernelGGL(sssp_small, dim3(1), dim3(SMALL_BLK), 0, g->stream, a);
        for (int v = 0; v < a.fa.nviews; ++v)
            hipLaunchKernelGGL(sssp_level_a, dim3(G), dim3(BLK), 0, g->stream, a, v);
        hipLaunchKernelGGL(sssp_level_b, dim3(G), dim3(BLK), 0, g->stream, a);
        MG_HIP(hipGetLastError());
        return MG_OK;
    };
    // Bellman-Ford with non-negative weights settles within n rounds.
    const int max_steps = (int)std::min<int64_t>((int64_t)o.iter_num, (int64_t)g->n + 1) + 2;
    if ((rc = drive(g, step, max_steps))) return rc;
    return finish_search(g, stats, ws.vals, INT_MAX, dist_out, ws.vals, 4 * (size_t)g->n, o.out_on_device, timer);
}
