// Single-source shortest paths: the engine's specialised path for the sssp program
// (Algorithms/SSSP/sssp.h).
//
// The reference program is push-style frontier Bellman-Ford: expand emits
// dist[v] + w for out-edges that improve on dist[nbr] (sssp.h:342-357), contract
// atomicMin's it into gather_tmp (sssp.h:384-401), apply/post_apply commit the
// minimum and mark changed vertices (sssp.h:242-280), and changed vertices form the
// next frontier.  Here one fused pass per level relaxes an edge straight into dist
// with atomicMin and marks the target in the next-frontier bitmap when it improved;
// the level loop, frontier rotation and termination stay on the device (same step
// structure as BFS: a single-workgroup kernel for small frontiers and the level
// close, two whole-GPU kernels per large level).  Integer shortest distances are
// unique, so the result equals the reference's: dist[src] = 0, unreached = INT_MAX
// (sssp.h:42-43, 95-100).
#include <climits>

#include "frontier.h"

namespace mg {

struct SsspArgs {
    FrontierArgs fa;
    int *dist;
    const int *w;
    int G;
};

struct SsspPolicy {
    const SsspArgs &a;
    long long nf = 0, mf = 0, ex = 0;
    __device__ explicit SsspPolicy(const SsspArgs &args) : a(args) {}
    __device__ int load(int v) const { return a.dist[v]; }
    __device__ bool visit(int view, int, int d, int slot, int u, bool valid, int L) {
        bool newly = false;
        if (valid) {
            ++ex;
            const long long nd = (long long)d + a.w[edge_id(a.fa, view, slot)];
            if (nd < (long long)a.dist[u] && (int)nd < atomicMin(&a.dist[u], (int)nd) && mark_next(a.fa, L, u)) {
                newly = true;
                ++nf;
                mf += view_degree(a.fa, u);
            }
        }
        return newly;
    }
};

__global__ void sssp_seed(SsspArgs a, int src, int iter_limit) {
    Ctl *c = a.fa.ctl;
    a.dist[src] = 0;
    a.fa.bm[0][src >> 5] = 1u << (src & 31);
    a.fa.q[0][0] = src;
    ctl_reset(c, a.fa.n, iter_limit, 1);
    c->mf_cur = view_degree(a.fa, src);
}

__global__ __launch_bounds__(SMALL_BLK) void sssp_small(SsspArgs a) {
    Ctl *c = a.fa.ctl;
    const FrontierArgs &fa = a.fa;
    __shared__ int s_go, s_L, s_qv, s_fin, s_cnt;
    if (c->pending) {
        const int L = c->level;
        const int na = fa.nviews * a.G;
        long long nf = 0, mf = 0, ex = 0;
        for (int i = threadIdx.x; i < na + a.G; i += SMALL_BLK) {
            const LevelPart p = fa.part[i < na ? i : 2 * a.G + (i - na)];
            nf += p.nf;
            mf += p.mf;
            ex += p.ex;
        }
        nf = block_sum<SMALL_BLK>(nf);
        mf = block_sum<SMALL_BLK>(mf);
        ex = block_sum<SMALL_BLK>(ex);
        if (threadIdx.x == 0) {
            c->edges_examined += ex;
            advance_level(c, L, nf, mf);
            c->pending = 0;
            c->qvalid = 0;
        }
        __syncthreads();
    }
    for (;;) {
        if (threadIdx.x == 0) {
            const int L = c->level;
            s_go = !c->done && c->qsize[L % 3] <= SMALL_Q && c->mf_cur <= SMALL_E;
            s_L = L;
            s_qv = c->qvalid;
            s_fin = c->qsize[L % 3];
            s_cnt = 0;
        }
        __syncthreads();
        if (!s_go) break;
        const int L = s_L, cur = L % 3, prv = (L + 2) % 3;
        if (!(s_qv & (1 << cur))) {
            queue_from_bitmap<SMALL_BLK>(fa, L, &s_cnt);
            if (threadIdx.x == 0) {
                s_fin = s_cnt;
                s_cnt = 0;
            }
        }
        if (s_qv & (1 << prv)) {
            const int *qp = fa.q[prv];
            for (int i = threadIdx.x; i < c->qsize[prv]; i += SMALL_BLK) fa.bm[prv][qp[i] >> 5] = 0;
        } else {
            clear_prev_sweep(fa, L, threadIdx.x, SMALL_BLK);
        }
        __syncthreads();
        SsspPolicy pol(a);
        int *qout = fa.q[(L + 1) % 3];
        auto emit = [&](int u, bool newly) { wave_append(newly, u, qout, &s_cnt); };
        const int *qin = fa.q[cur];
        for (int view = 0; view < fa.nviews; ++view)
            for (int t0 = 0; t0 < s_fin; t0 += SMALL_BLK) {
                const int i = t0 + (int)threadIdx.x;
                expand_tile<SMALL_BLK, false>(fa, pol, view, L, i < s_fin ? qin[i] : -1, emit);
            }
        const long long mf = block_sum<SMALL_BLK>(pol.mf);
        const long long ex = block_sum<SMALL_BLK>(pol.ex);
        if (threadIdx.x == 0) {
            c->small_levels++;
            c->edges_examined += ex;
            advance_level(c, L, s_cnt, mf);
            c->qvalid = (1 << cur) | (1 << ((L + 1) % 3));
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(BLK) void sssp_level_a(SsspArgs a, int view) {
    Ctl *c = a.fa.ctl;
    if (c->done) return;
    const int L = c->level;
    if (view == 0) clear_prev_sweep(a.fa, L, blockIdx.x * BLK + threadIdx.x, gridDim.x * BLK);
    SsspPolicy pol(a);
    expand_bitmap<BLK>(a.fa, pol, view, L, a.fa.bm[L % 3]);
    write_part<BLK>(a.fa, view * a.G + blockIdx.x, pol.nf, pol.mf, pol.ex);
}

__global__ __launch_bounds__(BLK) void sssp_level_b(SsspArgs a) {
    Ctl *c = a.fa.ctl;
    if (c->done) return;
    SsspPolicy pol(a);
    expand_chunks<BLK>(a.fa, pol, c->level);
    write_part<BLK>(a.fa, 2 * a.G + blockIdx.x, pol.nf, pol.mf, pol.ex);
    if (blockIdx.x == 0 && threadIdx.x == 0) c->pending = 1;
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
    a.G = grid_blocks(std::max<int64_t>(g->n / 4, g->m / 64), BLK);

    MG_HIP(hipEventRecord(g->ev0, g->stream));
    if ((rc = launch_fill(g, ws.vals, INT_MAX, g->n))) return rc;
    for (int i = 0; i < 3; ++i)
        if ((rc = launch_fill(g, ws.bm[i], 0u, ws.nwords))) return rc;
    hipLaunchKernelGGL(sssp_seed, dim3(1), dim3(1), 0, g->stream, a, src, o.iter_num);
    MG_HIP(hipGetLastError());
    auto step = [&]() -> int {
        hipLaunchKernelGGL(sssp_small, dim3(1), dim3(SMALL_BLK), 0, g->stream, a);
        for (int v = 0; v < a.fa.nviews; ++v)
            hipLaunchKernelGGL(sssp_level_a, dim3(a.G), dim3(BLK), 0, g->stream, a, v);
        hipLaunchKernelGGL(sssp_level_b, dim3(a.G), dim3(BLK), 0, g->stream, a);
        MG_HIP(hipGetLastError());
        return MG_OK;
    };
    // Bellman-Ford with non-negative weights settles within n rounds.
    const int64_t max_steps = std::min<int64_t>((int64_t)o.iter_num, (int64_t)g->n + 1) + 3;
    if ((rc = drive(g, step, max_steps))) return rc;
    return finish_search(g, stats, ws.vals, INT_MAX, dist_out, ws.vals, 4 * (size_t)g->n, o.out_on_device, timer);
}
