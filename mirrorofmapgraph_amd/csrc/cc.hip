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
// Here the same labels come from union-find, where every union hooks the larger
// root under the smaller one with an atomic compare-and-swap.  Parents only ever
// point to smaller ids, so a component's minimum vertex is never hooked under
// anything and ends as its root: label[v] = root(v) = min id of the component.
//
// Symmetric graphs use the subgraph-sampling scheme of Afforest (Sutton, Ben-Nun
// and Barak, IPDPS 2018):
//   1. unite every vertex with its first two neighbours, compress;
//   2. sample vertices and take the most frequent root as the large component C;
//   3. unite the remaining edges of every vertex outside C (an edge between C and a
//      vertex outside C is seen from the outside end, adjacency being symmetric);
//   4. compress: label[v] = root(v).
// Other graphs unite every CSR edge, which gives the weak components.
#include "frontier.h"

namespace mg {

constexpr int CC_SAMPLE_ROUNDS = 2;
constexpr int CC_SAMPLES = 1024;
// Step 1 takes the plain-store first pass (cc_link_first) on graphs of mean degree at least this, and
// union-find rounds (cc_link_round) below it.  The first pass points every vertex at its smallest neighbour,
// which is shallow where the smallest neighbour is mostly a hub (R-MAT scale 22, mean degree 30.6: CC 0.238 ->
// 0.141 ms) but follows the mesh on grids (chains up to the diameter: 2048^2 grid, mean degree 3.8, 0.752 ->
// 1.080 ms), whose compress then walks them; union-find rounds halve those paths as they go.
constexpr int CC_PLAIN_FIRST_DEGREE = 8;

// Parent loads may return a stale value (the caches of this CU/XCD): parents only
// ever move to ancestors, so a stale parent is still an ancestor, a stale "root" is
// caught by the hooking compare-and-swap (device scope), and two vertices seen under
// one root really are in one component.  The near caches are cheaper than the
// device-coherent path, which costs more than the extra retries.
__device__ __forceinline__ int ld_parent(const int *p, int x) {
    return __hip_atomic_load(&p[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
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
    bool half;  // unite each undirected edge once (u < v)
    __device__ int load(int) const { return 0; }
    __device__ bool visit(int, int v, int, int, int u, bool valid, int) {
        if (valid && (!half || u < v)) uf_unite(p, v, u);
        return false;
    }
};

__global__ void cc_init(FrontierArgs fa, int *p) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < fa.n; v += gridDim.x * blockDim.x) p[v] = v;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl_reset(fa.ctl, fa.n, 1, fa.n);
}

// Afforest step 1, first neighbour, with the parents' initialisation: parent[v] =
// min(v, first neighbour of v).  Only the thread of v writes parent[v] here, and every
// parent points to a smaller id or to itself, so the pointers form a forest whose
// roots are minima without any compare-and-swap.  A first neighbour larger than v
// is united in the second pass.
__global__ void cc_link_first(FrontierArgs fa, int *p) {
    const int *off = fa.off[0], *adj = fa.adj[0];
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < fa.n; v += gridDim.x * blockDim.x) {
        const int e = off[v];
        p[v] = e < off[v + 1] ? min(v, adj[e]) : v;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl_reset(fa.ctl, fa.n, 1, fa.n);
}

// Afforest step 1, second pass: the first neighbour where it is larger than v, and
// the second neighbour, by union-find.
__global__ void cc_link_second(FrontierArgs fa, int *p) {
    const int *off = fa.off[0], *adj = fa.adj[0];
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < fa.n; v += gridDim.x * blockDim.x) {
        const int e = off[v], end = off[v + 1];
        if (e < end) {
            const int u0 = adj[e];
            if (u0 > v) uf_unite(p, v, u0);
            if (e + 1 < end) uf_unite(p, v, adj[e + 1]);
        }
    }
}

// Afforest step 1 by union-find alone: unite v with its r-th neighbour (after cc_init).
__global__ void cc_link_round(FrontierArgs fa, int *p, int r) {
    const int *off = fa.off[0], *adj = fa.adj[0];
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < fa.n; v += gridDim.x * blockDim.x) {
        const int e = off[v] + r;
        if (e < off[v + 1]) uf_unite(p, v, adj[e]);
    }
}

__global__ void cc_compress(FrontierArgs fa, int *p) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < fa.n; v += gridDim.x * blockDim.x) {
        int r = v;
        for (int pr = p[r]; pr != r; pr = p[r]) r = pr;
        p[v] = r;
    }
}

// Afforest step 2: most frequent root among CC_SAMPLES hashed vertices (one
// workgroup; roots counted in an LDS hash table, then the largest count, ties to the
// smaller root).  Which component is skipped only decides the work of step 3, not the
// labels.
__global__ __launch_bounds__(CC_SAMPLES) void cc_sample(FrontierArgs fa, const int *p, uint64_t seed) {
    constexpr int H = 2 * CC_SAMPLES;
    __shared__ int s_key[H];
    __shared__ int s_cnt[H];
    __shared__ unsigned long long s_best[CC_SAMPLES / WAVE];
    const int t = threadIdx.x;
    for (int i = t; i < H; i += CC_SAMPLES) {
        s_key[i] = -1;
        s_cnt[i] = 0;
    }
    uint64_t h = (seed + (uint64_t)t * 0x9e3779b97f4a7c15ULL);
    h = (h ^ (h >> 30)) * 0xbf58476d1ce4e5b9ULL;
    h = (h ^ (h >> 27)) * 0x94d049bb133111ebULL;
    h ^= h >> 31;
    const int r = p[(int)(h % (uint64_t)fa.n)];
    __syncthreads();
    // insert r (linear probing; at most CC_SAMPLES distinct keys in 2 * CC_SAMPLES slots)
    uint32_t slot = ((uint32_t)r * 2654435761u) % H;
    for (;;) {
        const int old = atomicCAS(&s_key[slot], -1, r);
        if (old == -1 || old == r) break;
        slot = (slot + 1) % H;
    }
    atomicAdd(&s_cnt[slot], 1);
    __syncthreads();
    // the largest (count, -root) over the slots
    unsigned long long best = 0;
    for (int i = t; i < H; i += CC_SAMPLES)
        if (s_cnt[i] > 0) {
            const unsigned long long k = ((unsigned long long)s_cnt[i] << 32) | (0xFFFFFFFFu - (uint32_t)s_key[i]);
            best = k > best ? k : best;
        }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long y = __shfl_xor(best, o);
        best = y > best ? y : best;
    }
    if ((t & (WAVE - 1)) == 0) s_best[t / WAVE] = best;
    __syncthreads();
    if (t == 0) {
        for (int w = 1; w < CC_SAMPLES / WAVE; ++w) best = s_best[w] > best ? s_best[w] : best;
        fa.ctl->aux = (int)(0xFFFFFFFFu - (uint32_t)(best & 0xFFFFFFFFu));  // the skipped component's root
        fa.ctl->nf_prev = (int)(best >> 32);
    }
}

// Afforest step 3: remaining edges of every vertex outside the sampled component.
__global__ __launch_bounds__(BLK) void cc_finish(FrontierArgs fa, int *p) {
    CcPolicy pol{p, false};
    auto no_emit = [](int, bool) {};
    const int skip = fa.ctl->aux;
    const int *off = fa.off[0];
    for (int base = blockIdx.x * BLK; base < fa.n; base += gridDim.x * BLK) {
        const int v = base + (int)threadIdx.x;
        int beg = 0, deg = 0;
        if (v < fa.n && p[v] != skip) {
            beg = off[v] + CC_SAMPLE_ROUNDS;
            deg = max(0, off[v + 1] - beg);
        }
        expand_tile_range<BLK, true>(fa, pol, 0, 0, v < fa.n ? v : -1, beg, deg, no_emit);
    }
}

// Every CSR edge (general graphs): tiles through LDS, hubs as edge chunks.
__global__ __launch_bounds__(BLK) void cc_hook(FrontierArgs fa, int *p) {
    CcPolicy pol{p, false};
    auto no_emit = [](int, bool) {};
    for (int base = blockIdx.x * BLK; base < fa.n; base += gridDim.x * BLK) {
        const int v = base + (int)threadIdx.x;
        expand_tile<BLK, true>(fa, pol, 0, 0, v < fa.n ? v : -1, no_emit);
    }
}

__global__ __launch_bounds__(BLK) void cc_hook_chunks(FrontierArgs fa, int *p) {
    CcPolicy pol{p, false};
    expand_chunks<BLK>(fa, pol, 0);
}

__global__ void cc_done(Ctl *c) {
    c->level = 1;
    c->done = 1;
}

}  // namespace mg

using namespace mg;

extern "C" int mg_cc(mg_graph *g, int32_t *label_out, const mg_options *opt, mg_stats *stats) {
    CallTimer timer;
    mg_options o;
    int rc = begin_search(g, opt, &o);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g->mu);
    if (o.engine == MG_ENGINE_GAS) {
        // the reference runs cc with directed = false (cc.cu:289)
        mg_options og = o;
        og.directed = (g->flags & (MG_GRAPH_SYMMETRIC | MG_GRAPH_BUILD_CSC)) ? 0 : 1;
        return gas_cc(g, label_out, og, stats, timer);
    }
    Workspace &ws = g->ws;
    FrontierArgs fa = make_frontier_args(g, 1);
    int *p = ws.vals;
    const int G = grid_blocks_cap(std::max<int64_t>(g->n / 4, g->m / 64), BLK, BIG_GRID);
    const int Gn = grid_blocks_cap(g->n, BLK, BIG_GRID);
    MG_HIP(hipEventRecord(g->ev0, g->stream));
    if ((g->flags & MG_GRAPH_SYMMETRIC) && o.engine != MG_ENGINE_GAS && g->n > 0) {
        if (g->m >= (int64_t)CC_PLAIN_FIRST_DEGREE * g->n) {
            hipLaunchKernelGGL(cc_link_first, dim3(Gn), dim3(BLK), 0, g->stream, fa, p);
            hipLaunchKernelGGL(cc_compress, dim3(Gn), dim3(BLK), 0, g->stream, fa, p);
            hipLaunchKernelGGL(cc_link_second, dim3(Gn), dim3(BLK), 0, g->stream, fa, p);
            hipLaunchKernelGGL(cc_compress, dim3(Gn), dim3(BLK), 0, g->stream, fa, p);
        } else {
            hipLaunchKernelGGL(cc_init, dim3(Gn), dim3(BLK), 0, g->stream, fa, p);
            for (int r = 0; r < CC_SAMPLE_ROUNDS; ++r) {
                hipLaunchKernelGGL(cc_link_round, dim3(Gn), dim3(BLK), 0, g->stream, fa, p, r);
                hipLaunchKernelGGL(cc_compress, dim3(Gn), dim3(BLK), 0, g->stream, fa, p);
            }
        }
        hipLaunchKernelGGL(cc_sample, dim3(1), dim3(CC_SAMPLES), 0, g->stream, fa, (const int *)p,
                           (uint64_t)0x5eed);
        hipLaunchKernelGGL(cc_finish, dim3(G), dim3(BLK), 0, g->stream, fa, p);
    } else {
        hipLaunchKernelGGL(cc_init, dim3(Gn), dim3(BLK), 0, g->stream, fa, p);
        hipLaunchKernelGGL(cc_hook, dim3(G), dim3(BLK), 0, g->stream, fa, p);
    }
    hipLaunchKernelGGL(cc_hook_chunks, dim3(G), dim3(BLK), 0, g->stream, fa, p);
    hipLaunchKernelGGL(cc_compress, dim3(Gn), dim3(BLK), 0, g->stream, fa, p);
    hipLaunchKernelGGL(cc_done, dim3(1), dim3(1), 0, g->stream, ws.ctl);
    MG_HIP(hipGetLastError());
    return finish_search(g, stats, ws.vals, -1, label_out, ws.vals, 4 * (size_t)g->n, o.out_on_device, timer);
}
