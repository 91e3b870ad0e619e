// PageRank: the engine's specialised path for the pagerank program
// (Algorithms/PageRank/PageRank.h).
//
// The reference program gathers over in-edges (GATHER_IN_EDGES, PageRank.h:197-200)
// the value rank[u] / outdeg[u] (gather_edge, :233-243), sums it (gather_sum), and
// applies new = 0.15f + (1.0f - 0.15f) * sum, marking the vertex changed when
// |old - new| >= 0.01f (apply, :257-283); the next frontier is the out-neighbours
// of changed vertices (expand_edge, :355-362), deduplicated.  Gathers of one
// iteration all read the previous ranks (the enactor runs gather before apply),
// i.e. frontier-restricted Jacobi.  The enactor materialises (dst, value) pairs
// per in-edge and reduces them with a segmented reduction (enactor:1503-1596).
//
// Here the gather is a pull over the in-edge view that fuses gather, apply and the
// frontier test: a vertex is active at iteration i > 0 exactly when one of its
// in-neighbours changed at iteration i-1 (that is the reference's expand), so one
// pass per iteration reads each in-edge once, sums precomputed contributions
// contrib[u] = rank[u] / (float)outdeg[u] (the same float division as gather_edge),
// and applies.  Summation order is fixed, so results are deterministic:
//   in-degree <= PR_T1   one lane, sequentially in in-edge order;
//   <= PR_T2             one wave: lane l sums edges l, l+64, ... in order, then an
//                        xor-butterfly over offsets 32,16,8,4,2,1;
//   >  PR_T2             chunks of PR_T2 edges summed as above, then the chunk
//                        sums added in chunk order.
// MG_PR_ALL updates every vertex each iteration (all-vertex Jacobi) and can stop on
// the L1 change, reduced in a fixed order as well.
#include <vector>

#include "frontier.h"

namespace mg {

constexpr int PR_T1 = 64;
constexpr int PR_T2 = 1024;
constexpr int PR_MAX_GRID = 32768;  // workgroups of pr_k1 / pr_k2

struct PrArgs {
    Ctl *ctl;
    const int *in_off, *in_adj, *row_off;
    float *rank;
    float *contrib[2];
    uint32_t *chg[3];
    const int *mid;
    const ChunkRec *chunks;
    const int *bigv;
    float *part;
    int *pact;
    double *partials;    // per-workgroup L1 change
    LevelPart *cpart;    // per-workgroup count of changed vertices (nf)
    int n, nwords, nmid, nchunks, nbig, G1, G2, G3;  // G3: workgroups of pr_k3b (0: not run)
    int mode;
    float base, damp, eps;
    double tol_n;
};

struct PrIter {
    int it;
    const float *c_old;
    float *c_new;
    const uint32_t *prev;
    uint32_t *cur;
};

__device__ __forceinline__ PrIter pr_iter(const PrArgs &a) {
    PrIter s;
    s.it = a.ctl->level;
    s.c_old = a.contrib[s.it & 1];
    s.c_new = a.contrib[(s.it + 1) & 1];
    s.prev = a.chg[(s.it + 2) % 3];
    s.cur = a.chg[s.it % 3];
    return s;
}

// Applies the gathered sum s to v; returns |new - old| (0 when inactive).
__device__ __forceinline__ double pr_apply(const PrArgs &a, const PrIter &I, int v, float s, bool active,
                                           int &nchanged) {
    const int od = a.row_off[v + 1] - a.row_off[v];
    if (!active) {
        I.c_new[v] = I.c_old[v];
        return 0.0;
    }
    const float old = a.rank[v];
    const float nv = a.base + a.damp * s;
    a.rank[v] = nv;
    I.c_new[v] = nv / (float)od;
    if (a.mode == MG_PR_GAS && !(fabsf(old - nv) < a.eps)) {
        atomicOr(&I.cur[v >> 5], 1u << (v & 31));
        if (od > 0) ++nchanged;  // only vertices with out-edges put anything in the next frontier
    }
    return fabs((double)nv - (double)old);
}

// Stores a workgroup's L1 change and changed-vertex count in slot `slot`.
template <int BLOCK>
__device__ __forceinline__ void pr_write_part(const PrArgs &a, int slot, double delta, long long nchanged) {
    delta = block_sum<BLOCK>(delta);
    nchanged = block_sum<BLOCK>(nchanged);
    if (threadIdx.x == 0) {
        a.partials[slot] = delta;
        a.cpart[slot].nf = nchanged;
    }
}

__global__ void pr_init(PrArgs a, int iter_limit) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < a.n; v += gridDim.x * blockDim.x) {
        a.rank[v] = a.base;
        a.contrib[0][v] = a.base / (float)(a.row_off[v + 1] - a.row_off[v]);
        a.contrib[1][v] = a.contrib[0][v];
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.nwords; i += gridDim.x * blockDim.x) {
        a.chg[0][i] = 0;
        a.chg[1][i] = 0;
        a.chg[2][i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl_reset(a.ctl, a.n, iter_limit, a.n);
}

// Low in-degree vertices, one lane each, in vertex order.
__global__ __launch_bounds__(BLK) void pr_k1(PrArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const PrIter I = pr_iter(a);
    const bool all = a.mode == MG_PR_ALL || I.it == 0;
    // the bitmap filled at iteration it+1 is cleared now
    uint32_t *clr = a.chg[(I.it + 1) % 3];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.nwords; i += gridDim.x * blockDim.x) clr[i] = 0;
    double delta = 0;
    int nchanged = 0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < a.n; v += gridDim.x * blockDim.x) {
        const int beg = a.in_off[v], end = a.in_off[v + 1];
        if (end - beg > PR_T1) continue;
        float s = 0.0f;
        bool act = all;
        // four in-edges in flight (full groups load unconditionally), summed in order
        int e = beg;
        for (; e + 4 <= end; e += 4) {
            int u[4];
            float cv[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) u[j] = a.in_adj[e + j];
#pragma unroll
            for (int j = 0; j < 4; ++j) cv[j] = I.c_old[u[j]];
            if (!all) {
#pragma unroll
                for (int j = 0; j < 4; ++j) act = act || test_bit(I.prev, u[j]);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) s = s + cv[j];
        }
        for (; e < end; ++e) {
            const int u = a.in_adj[e];
            if (!all) act = act || test_bit(I.prev, u);
            s = s + I.c_old[u];
        }
        delta += pr_apply(a, I, v, s, act, nchanged);
    }
    pr_write_part<BLK>(a, blockIdx.x, delta, nchanged);
}

// Wave sum of edges [beg, end): lane-strided, then xor-butterfly.  `act` is OR-ed
// over the wave as well.
__device__ __forceinline__ float wave_gather(const PrArgs &a, const PrIter &I, int beg, int end, bool &act) {
    const int lane = lane_id();
    const bool all = act;  // activity already known: skip the bitmap
    float s = 0.0f;
    bool any = false;
    // lane l sums edges l, l+64, ... in order, four in flight (full groups load
    // unconditionally), then the remainder one at a time
    int e = beg + lane;
    for (; e + 3 * WAVE < end; e += 4 * WAVE) {
        int u[4];
        float cv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) u[j] = a.in_adj[e + j * WAVE];
#pragma unroll
        for (int j = 0; j < 4; ++j) cv[j] = I.c_old[u[j]];
        if (!all) {
#pragma unroll
            for (int j = 0; j < 4; ++j) any = any || test_bit(I.prev, u[j]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) s = s + cv[j];
    }
    for (; e < end; e += WAVE) {
        const int u = a.in_adj[e];
        if (!all) any = any || test_bit(I.prev, u);
        s = s + I.c_old[u];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s = s + __shfl_xor(s, o);
    act = act || (wave_ballot(any) != 0);
    return s;
}

// Medium vertices (a wave each) and chunks of high in-degree vertices.
__global__ __launch_bounds__(BLK) void pr_k2(PrArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const PrIter I = pr_iter(a);
    const bool all = a.mode == MG_PR_ALL || I.it == 0;
    const int lane = lane_id();
    const int wave = (blockIdx.x * BLK + threadIdx.x) / WAVE, nwaves = gridDim.x * BLK / WAVE;
    double delta = 0;
    int nchanged = 0;
    for (int t = wave; t < a.nmid + a.nchunks; t += nwaves) {
        bool act = all;
        if (t < a.nmid) {
            const int v = a.mid[t];
            const float s = wave_gather(a, I, a.in_off[v], a.in_off[v + 1], act);
            if (lane == 0) delta += pr_apply(a, I, v, s, act, nchanged);
        } else {
            const int ci = t - a.nmid;
            const ChunkRec r = a.chunks[ci];
            const float s = wave_gather(a, I, r.beg, r.end, act);
            if (lane == 0) {
                a.part[ci] = s;
                a.pact[ci] = act;
            }
        }
    }
    pr_write_part<BLK>(a, a.G1 + blockIdx.x, delta, nchanged);
}

// High in-degree vertices, a thread each: their chunk sums added in chunk order
// (eight loads in flight), then applied.  Partials go to slots [G1 + G2, + G3).
__global__ __launch_bounds__(BLK) void pr_k3b(PrArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const PrIter I = pr_iter(a);
    const bool all = a.mode == MG_PR_ALL || I.it == 0;
    double delta = 0;
    int nchanged = 0;
    for (int k = blockIdx.x * BLK + (int)threadIdx.x; k < a.nbig; k += gridDim.x * BLK) {
        const int c0 = a.bigv[k], c1 = a.bigv[k + 1];
        const int v = a.chunks[c0].v;
        float s = a.part[c0];
        bool act = a.pact[c0] != 0;
        int ci = c0 + 1;
        for (; ci + 8 <= c1; ci += 8) {
            float p[8];
            int q[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                p[j] = a.part[ci + j];
                q[j] = a.pact[ci + j];
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                s = s + p[j];
                act = act || q[j] != 0;
            }
        }
        for (; ci < c1; ++ci) {
            s = s + a.part[ci];
            act = act || a.pact[ci] != 0;
        }
        delta += pr_apply(a, I, v, s, act || all, nchanged);
    }
    pr_write_part<BLK>(a, a.G1 + a.G2 + blockIdx.x, delta, nchanged);
}

// The iteration close in two stages, each in a fixed order (the same bits every run).
// Tens of thousands of per-workgroup partials read by one workgroup took 22 us per
// iteration at scale 22; first PR_RED workgroups sum fixed ranges of them into slots
// np .. np + PR_RED - 1, then one workgroup sums those.
constexpr int PR_RED = 64;
__global__ __launch_bounds__(BLK) void pr_k3a(PrArgs a) {
    if (a.ctl->done) return;
    const int np = a.G1 + a.G2 + a.G3;
    const int per = (np + PR_RED - 1) / PR_RED;
    const int lo = blockIdx.x * per, hi = min(np, lo + per);
    double delta = 0;
    long long nch = 0;
    for (int i = lo + (int)threadIdx.x; i < hi; i += BLK) {
        delta += a.partials[i];
        nch += a.cpart[i].nf;
    }
    delta = block_sum<BLK>(delta);
    nch = block_sum<BLK>(nch);
    if (threadIdx.x == 0) {
        a.partials[np + blockIdx.x] = delta;
        a.cpart[np + blockIdx.x].nf = nch;
    }
}

// One workgroup: closes the iteration from the PR_RED range sums.
__global__ __launch_bounds__(SMALL_BLK) void pr_k3(PrArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const PrIter I = pr_iter(a);
    double delta = 0;
    long long nch = 0;
    const int np = a.G1 + a.G2 + a.G3;
    for (int i = threadIdx.x; i < PR_RED; i += SMALL_BLK) {
        delta += a.partials[np + i];
        nch += a.cpart[np + i].nf;
    }
    const double l1 = block_sum<SMALL_BLK>(delta);
    nch = block_sum<SMALL_BLK>(nch);
    if (threadIdx.x == 0) {
        c->dsum = l1;
        c->level = I.it + 1;
        c->changed = (int)nch;
        bool done = I.it + 1 >= c->iter_limit;
        if (a.mode == MG_PR_GAS && nch == 0) done = true;
        if (a.mode == MG_PR_ALL && a.tol_n > 0 && l1 < a.tol_n) done = true;
        c->done = done;
    }
}

// Builds (once per graph and in-edge view) the medium-vertex list and the chunk
// records of high in-degree vertices, on the host from the in-edge offsets.
static int pr_buckets(mg_graph *g, const int *d_in_off) {
    Workspace &ws = g->ws;
    if (ws.pr_view == d_in_off) return MG_OK;
    const int n = g->n;
    std::vector<int32_t> off((size_t)n + 1);
    MG_HIP(copy_on(g->stream, off.data(), d_in_off, 4 * ((size_t)n + 1), hipMemcpyDeviceToHost));
    std::vector<int32_t> mid, bigv;
    std::vector<ChunkRec> ch;
    for (int v = 0; v < n; ++v) {
        const int d = off[v + 1] - off[v];
        if (d <= PR_T1) continue;
        if (d <= PR_T2) {
            mid.push_back(v);
        } else {
            bigv.push_back((int)ch.size());
            for (int b = off[v]; b < off[v + 1]; b += PR_T2) ch.push_back(ChunkRec{v, b, std::min(b + PR_T2, off[v + 1]), 0});
        }
    }
    bigv.push_back((int)ch.size());
    free_dev(ws.pr_mid);
    free_dev(ws.pr_chunks);
    free_dev(ws.pr_bigv);
    free_dev(ws.pr_part);
    free_dev(ws.pr_pact);
    ws.pr_mid = nullptr;
    ws.pr_chunks = nullptr;
    ws.pr_bigv = nullptr;
    ws.pr_part = nullptr;
    ws.pr_pact = nullptr;
    ws.pr_view = nullptr;
    int rc;
    if ((rc = alloc_i32(&ws.pr_mid, mid.size())) || (rc = alloc_chunks(&ws.pr_chunks, ch.size())) ||
        (rc = alloc_i32(&ws.pr_bigv, bigv.size())) || (rc = alloc_f32(&ws.pr_part, ch.size())) ||
        (rc = alloc_i32(&ws.pr_pact, ch.size())))
        return rc;
    if (!mid.empty()) MG_HIP(copy_on(g->stream, ws.pr_mid, mid.data(), 4 * mid.size(), hipMemcpyHostToDevice));
    if (!ch.empty()) MG_HIP(copy_on(g->stream, ws.pr_chunks, ch.data(), sizeof(ChunkRec) * ch.size(), hipMemcpyHostToDevice));
    MG_HIP(copy_on(g->stream, ws.pr_bigv, bigv.data(), 4 * bigv.size(), hipMemcpyHostToDevice));
    ws.pr_nmid = (int)mid.size();
    ws.pr_nchunks = (int)ch.size();
    ws.pr_nbig = (int)bigv.size() - 1;
    ws.pr_view = d_in_off;
    return MG_OK;
}

}  // namespace mg

using namespace mg;

extern "C" int mg_pagerank(mg_graph *g, float *rank_out, const mg_options *opt, mg_stats *stats) {
    CallTimer timer;
    mg_options o;
    int rc = begin_search(g, opt, &o);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g->mu);
    if (o.pr_mode != MG_PR_GAS && o.pr_mode != MG_PR_ALL)
        return set_error(MG_ERR_INVALID_ARG, "mg_pagerank: unknown pr_mode %d", o.pr_mode);
    const bool sym = g->flags & MG_GRAPH_SYMMETRIC;
    if (!sym && !g->col_off)
        return set_error(MG_ERR_UNSUPPORTED, "mg_pagerank: needs the in-edge view (MG_GRAPH_BUILD_CSC) or a symmetric graph");
    if (o.directed == 0 && !sym)
        return set_error(MG_ERR_UNSUPPORTED, "mg_pagerank: directed=0 is supported on symmetric graphs only");
    if (o.engine == MG_ENGINE_GAS) {
        if (o.pr_mode != MG_PR_GAS || o.pr_teleport != 0.15f || o.pr_active_eps != 0.01f)
            return set_error(MG_ERR_UNSUPPORTED, "MG_ENGINE_GAS runs the reference pagerank program (pr_mode "
                                                 "MG_PR_GAS, teleport 0.15, eps 0.01)");
        return gas_pagerank(g, rank_out, o, stats, timer);
    }
    g = search_graph(g, o, nullptr);  // the degree-ordered twin, when the graph keeps one
    if ((rc = ensure_workspace(g)) || (rc = ensure_pr_workspace(g))) return rc;
    const int *in_off = sym ? g->row_off : g->col_off;
    const int *in_adj = sym ? g->col : g->row_idx;
    if ((rc = pr_buckets(g, in_off))) return rc;
    Workspace &ws = g->ws;
    PrArgs a{};
    a.ctl = ws.ctl;
    a.in_off = in_off;
    a.in_adj = in_adj;
    a.row_off = g->row_off;
    a.rank = ws.rank;
    a.contrib[0] = ws.contrib[0];
    a.contrib[1] = ws.contrib[1];
    for (int i = 0; i < 3; ++i) a.chg[i] = ws.bm[i];
    a.mid = ws.pr_mid;
    a.chunks = ws.pr_chunks;
    a.bigv = ws.pr_bigv;
    a.part = ws.pr_part;
    a.pact = ws.pr_pact;
    a.partials = ws.partials;
    a.cpart = ws.part;
    a.n = g->n;
    a.nwords = ws.nwords;
    a.nmid = ws.pr_nmid;
    a.nchunks = ws.pr_nchunks;
    a.nbig = ws.pr_nbig;
    a.G1 = grid_blocks(g->n, BLK);
    a.G2 = grid_blocks((int64_t)(a.nmid + a.nchunks) * WAVE, BLK);
    a.mode = o.pr_mode;
    a.base = o.pr_teleport;
    a.damp = 1.0f - o.pr_teleport;
    a.eps = o.pr_active_eps;
    a.tol_n = (double)o.pr_tol * (double)g->n;

    // high in-degree vertices' chunk sums
    a.G3 = a.nbig > 0 ? grid_blocks(a.nbig, BLK) : 0;
    // The gather is latency-bound: with a grid-stride loop each thread walks several
    // vertices (pr_k1) or each wave several lists (pr_k2) one after another.  Up to
    // PR_MAX_GRID workgroups give every vertex / list its own thread / wave and let the
    // hardware overlap them; the partial slots grow to match.
    auto blocks = [](int64_t items) { return (int)std::max<int64_t>(1, std::min<int64_t>((items + BLK - 1) / BLK, PR_MAX_GRID)); };
    a.G1 = blocks(g->n);
    a.G2 = blocks((int64_t)(a.nmid + a.nchunks) * WAVE);
    const int np = a.G1 + a.G2 + a.G3;
    if (ws.pr_nslots < np + PR_RED) {  // the workgroups' slots, then the range sums
        free_dev(ws.pr_partials);
        free_dev(ws.pr_cpart);
        ws.pr_partials = nullptr;
        ws.pr_cpart = nullptr;
        ws.pr_nslots = 0;
        if ((rc = alloc_f64(&ws.pr_partials, np + PR_RED)) || (rc = alloc_parts(&ws.pr_cpart, np + PR_RED))) return rc;
        ws.pr_nslots = np + PR_RED;
    }
    a.partials = ws.pr_partials;
    a.cpart = ws.pr_cpart;

    MG_HIP(hipEventRecord(g->ev0, g->stream));
    hipLaunchKernelGGL(pr_init, dim3(grid_blocks(g->n, BLK)), dim3(BLK), 0, g->stream, a, o.iter_num);
    MG_HIP(hipGetLastError());
    auto step = [&]() -> int {
        hipLaunchKernelGGL(pr_k1, dim3(a.G1), dim3(BLK), 0, g->stream, a);
        hipLaunchKernelGGL(pr_k2, dim3(a.G2), dim3(BLK), 0, g->stream, a);
        if (a.G3) hipLaunchKernelGGL(pr_k3b, dim3(a.G3), dim3(BLK), 0, g->stream, a);
        hipLaunchKernelGGL(pr_k3a, dim3(PR_RED), dim3(BLK), 0, g->stream, a);
        hipLaunchKernelGGL(pr_k3, dim3(1), dim3(SMALL_BLK), 0, g->stream, a);
        MG_HIP(hipGetLastError());
        return MG_OK;
    };
    // the captured steps bake in the per-call constants: key on all of them
    uint64_t key = (uint64_t)a.mode;
    for (uint64_t x : {(uint64_t)__builtin_bit_cast(uint32_t, a.base), (uint64_t)__builtin_bit_cast(uint32_t, a.eps),
                       (uint64_t)__builtin_bit_cast(uint64_t, a.tol_n), (uint64_t)(uintptr_t)in_off})
        key = mix64(key ^ x);
    if ((rc = drive(g, step, (int64_t)std::max(o.iter_num, 0) + 1, KIND_PR, key))) return rc;
    rc = finish_search(g, stats, nullptr, 0, rank_out, ws.rank, 4 * (size_t)g->n, o.out_on_device, timer);
    if (rc) return rc;
    if (stats) {
        stats->edges_traversed = g->m * (int64_t)stats->iterations;
        stats->vertices_reached = g->n;
    }
    return MG_OK;
}
