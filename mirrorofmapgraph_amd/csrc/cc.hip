// Connected components: the engine's specialised path for the cc program
// (Algorithms/CC/cc.h).
//
// The reference propagates min labels over the frontier: label[v] = v, expand
// sends label[v] to neighbours whose label is larger (cc.h:289-301), contract
// atomicMin's into gather_tmp (cc.h:328-341), apply keeps the minimum (cc.h:189-216),
// and the driver forces directed = false so both edge directions are used
// (cc.cu:289).  The result is, for every vertex, the smallest vertex id of its
// weakly connected component.
//
// Here the same labels come from one union-find pass (no level loop):
//   1. parent[v] = the first neighbour u < v when one exists, else v;
//   2. every edge unites its endpoints, always hooking the larger root under the
//      smaller one with an atomic compare-and-swap;
//   3. label[v] = root(v).
// Parents only ever point to smaller ids, so a component's minimum vertex is never
// hooked under anything and ends as its root: label[v] = min id of the component,
// the reference's answer.  Each undirected edge of a symmetric graph is united once
// (from its larger endpoint); a directed graph unites every CSR edge, which gives
// the weak components as the reference's two-direction sweep does.
#include "frontier.h"

namespace mg {

__device__ __forceinline__ int ld_parent(const int *p, int x) {
    return __hip_atomic_load(&p[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving (parents only move to ancestors, so concurrent
// halving keeps every parent[] entry pointing to an ancestor).
__device__ __forceinline__ int uf_find(int *p, int x) {
    for (;;) {
        const int px = ld_parent(p, x);
        if (px == x) return x;
        const int ppx = ld_parent(p, px);
        if (ppx != px) __hip_atomic_store(&p[x], ppx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = ppx;
    }
}

__device__ __forceinline__ void uf_unite(int *p, int a, int b) {
    int ra = uf_find(p, a), rb = uf_find(p, b);
    while (ra != rb) {
        if (ra < rb) {
            const int t = ra;
            ra = rb;
            rb = t;
        }
        // hook the larger root ra under the smaller root rb
        const int old = atomicCAS(&p[ra], ra, rb);
        if (old == ra) return;
        ra = uf_find(p, old);
        rb = uf_find(p, rb);
    }
}

struct CcPolicy {
    int *p;
    bool half;  // symmetric graph: unite each undirected edge once (u < v)
    __device__ int load(int) const { return 0; }
    __device__ bool visit(int, int v, int, int, int u, bool valid, int) {
        if (valid && (!half || u < v)) uf_unite(p, v, u);
        return false;
    }
};

__global__ void cc_init(FrontierArgs fa, int *p) {
    const int *off = fa.off[0], *adj = fa.adj[0];
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < fa.n; v += gridDim.x * blockDim.x) {
        int par = v;
        const int end = min(off[v + 1], off[v] + 8);
        for (int e = off[v]; e < end; ++e)
            if (adj[e] < v) {
                par = adj[e];
                break;
            }
        p[v] = par;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl_reset(fa.ctl, fa.n, 1, fa.n);
}

// Every vertex's edges, tiled through LDS; hubs become edge chunks.
__global__ __launch_bounds__(BLK) void cc_hook(FrontierArgs fa, int *p, int half) {
    CcPolicy pol{p, half != 0};
    auto no_emit = [](int, bool) {};
    for (int base = blockIdx.x * BLK; base < fa.n; base += gridDim.x * BLK) {
        const int v = base + (int)threadIdx.x;
        expand_tile<BLK, true>(fa, pol, 0, 0, v < fa.n ? v : -1, no_emit);
    }
}

__global__ __launch_bounds__(BLK) void cc_hook_chunks(FrontierArgs fa, int *p, int half) {
    CcPolicy pol{p, half != 0};
    expand_chunks<BLK>(fa, pol, 0);
}

__global__ void cc_compress(FrontierArgs fa, int *p) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < fa.n; v += gridDim.x * blockDim.x) {
        int r = v;
        for (int pr = p[r]; pr != r; pr = p[r]) r = pr;
        p[v] = r;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        fa.ctl->level = 1;
        fa.ctl->done = 1;
    }
}

}  // namespace mg

using namespace mg;

extern "C" int mg_cc(mg_graph *g, int32_t *label_out, const mg_options *opt, mg_stats *stats) {
    CallTimer timer;
    mg_options o;
    int rc = begin_search(g, opt, &o);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g->mu);
    Workspace &ws = g->ws;
    FrontierArgs fa = make_frontier_args(g, 1);
    const int half = (g->flags & MG_GRAPH_SYMMETRIC) ? 1 : 0;
    const int G = grid_blocks(std::max<int64_t>(g->n / 4, g->m / 64), BLK);
    MG_HIP(hipEventRecord(g->ev0, g->stream));
    hipLaunchKernelGGL(cc_init, dim3(grid_blocks(g->n, BLK)), dim3(BLK), 0, g->stream, fa, ws.vals);
    hipLaunchKernelGGL(cc_hook, dim3(G), dim3(BLK), 0, g->stream, fa, ws.vals, half);
    hipLaunchKernelGGL(cc_hook_chunks, dim3(G), dim3(BLK), 0, g->stream, fa, ws.vals, half);
    hipLaunchKernelGGL(cc_compress, dim3(grid_blocks(g->n, BLK)), dim3(BLK), 0, g->stream, fa, ws.vals);
    MG_HIP(hipGetLastError());
    return finish_search(g, stats, ws.vals, -1, label_out, ws.vals, 4 * (size_t)g->n, o.out_on_device, timer);
}
