// Breadth-first search: the engine's specialised path for the bfs program
// (Algorithms/BFS/bfs.h; enactor loop GASengine/enactor_vertex_centric.cuh:2526-2872).
//
// The reference runs expand (expand_atomic, LDS/warp/CTA gathering) then contract
// (contract_atomic: label check-and-set, bitmask/hash culls) per level with a host
// round trip per level.  Here the level loop lives on the device:
//   * bfs_small    one 1024-thread workgroup runs whole levels back to back while
//                  the frontier is small (long-diameter graphs spend almost all
//                  levels here, with no launch per level);
//   * bfs_level_a  one level over the whole GPU: top-down (frontier queue, tiled
//                  through LDS, hubs split into edge chunks) or bottom-up (every
//                  unvisited vertex scans its in-edges for a parent in the frontier
//                  bitmap, which stays L2-resident: 512 KB at 4M vertices);
//   * bfs_level_b  the edge chunks / long bottom-up scans of the level, then the
//                  last workgroup closes the level (direction choice, queue
//                  rotation, termination) on the device.
// Depths are unique, so any traversal order yields the reference's labels
// (bfs.h:241-253: label = iteration + 1 on first visit; bfs.cu:56-143 on the CPU).
#include "frontier.h"

namespace mg {

struct BfsArgs {
    FrontierArgs fa;
    const int *in_off;   // bottom-up view (in-edges), nullptr when bottom-up is not allowed
    const int *in_adj;
    const int *row_off;  // CSR out-degree (direction heuristic)
    int *depth;
    uint32_t *visited;
    int *hard;
};

// Top-down visit: claim u through the visited bitmap, label it, mark it in the next
// frontier bitmap and append it to the next queue.
struct BfsPolicy {
    const BfsArgs &a;
    long long mf = 0;
    __device__ explicit BfsPolicy(const BfsArgs &args) : a(args) {}
    __device__ int load(int) const { return 0; }
    __device__ void visit(int, int, int, int, int u, bool valid, int L) {
        bool own = false;
        if (valid && !test_bit(a.visited, u)) own = set_bit_first(a.visited, u);
        if (own) {
            a.depth[u] = L + 1;
            atomicOr(&a.fa.bm[(L + 1) % 3][u >> 5], 1u << (u & 31));
            mf += view_degree(a.fa, u);
        }
        wave_append(own, u, a.fa.q[(L + 1) % 3], &a.fa.ctl->qsize[(L + 1) % 3]);
    }
};

// Closes level L (one thread): direction choice (Beamer, Asanovic, Patterson 2012:
// go bottom-up when the next frontier's edges exceed unexplored/alpha, back
// top-down when the frontier falls below n/beta), then queue rotation.
__device__ void bfs_level_end(const BfsArgs &a) {
    Ctl *c = a.fa.ctl;
    const int L = c->level;
    const long long mf = ld_dev(&c->mf_next);
    c->m_unexplored -= mf;
    int dir = c->dir;
    if (dir == DIR_BU) c->bu_levels++;
    const int nf = ld_dev(&c->qsize[(L + 1) % 3]);
    if (c->can_bu) {
        if (dir == DIR_TD && mf > c->m_unexplored / c->alpha) dir = DIR_BU;
        else if (dir == DIR_BU && (long long)nf * c->beta < c->n) dir = DIR_TD;
    }
    c->dir = dir;
    c->mf_cur = mf;
    rotate_level(c, L);
}

__global__ void bfs_init(BfsArgs a) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
    for (int i = tid; i < a.fa.n; i += nt) a.depth[i] = -1;
    for (int i = tid; i < a.fa.nwords; i += nt) {
        a.visited[i] = 0;
        a.fa.bm[0][i] = 0;
        a.fa.bm[1][i] = 0;
        a.fa.bm[2][i] = 0;
    }
}

__global__ void bfs_seed(BfsArgs a, int src, int iter_limit, int can_bu, int alpha, int beta, long long m_views) {
    Ctl *c = a.fa.ctl;
    const uint32_t bit = 1u << (src & 31);
    a.depth[src] = 0;
    a.visited[src >> 5] = bit;
    a.fa.bm[0][src >> 5] = bit;
    a.fa.q[0][0] = src;
    ctl_reset(c, a.fa.n, iter_limit, 1);
    const long long d = view_degree(a.fa, src);
    c->can_bu = can_bu;
    c->alpha = alpha > 0 ? alpha : 14;
    c->beta = beta > 0 ? beta : 24;
    c->m_unexplored = m_views - d;
    c->mf_cur = d;
}

// Single-workgroup levels while the frontier is small (top-down only).
__global__ __launch_bounds__(SMALL_BLK) void bfs_small(BfsArgs a) {
    Ctl *c = a.fa.ctl;
    __shared__ int s_go, s_L;
    __shared__ unsigned long long s_mf;
    for (;;) {
        if (threadIdx.x == 0) {
            const int L = c->level;
            s_go = !c->done && c->dir == DIR_TD && c->qsize[L % 3] <= SMALL_Q && c->mf_cur <= SMALL_E;
            s_L = L;
            s_mf = 0;
        }
        __syncthreads();
        if (!s_go) break;
        const int L = s_L;
        clear_prev<SMALL_BLK>(a.fa, L, 0, 1);
        __syncthreads();
        BfsPolicy pol(a);
        for (int view = 0; view < a.fa.nviews; ++view) td_tiles<SMALL_BLK, false>(a.fa, pol, view, L, 0, 1);
        if (pol.mf) atomicAdd(&s_mf, (unsigned long long)pol.mf);
        __syncthreads();
        if (threadIdx.x == 0) {
            c->mf_next = (long long)s_mf;
            c->small_levels++;
            bfs_level_end(a);
        }
        __syncthreads();
    }
}

// One level over the whole GPU: top-down for `view`, or the bottom-up sweep.
__global__ __launch_bounds__(BLK) void bfs_level_a(BfsArgs a, int view) {
    Ctl *c = a.fa.ctl;
    if (c->done) return;
    const int L = c->level;
    const int dir = c->dir;
    if (dir == DIR_BU && view != 0) return;
    if (view == 0) clear_prev<BLK>(a.fa, L, blockIdx.x, gridDim.x);
    long long mf = 0;
    if (dir == DIR_TD) {
        BfsPolicy pol(a);
        td_tiles<BLK, true>(a.fa, pol, view, L, blockIdx.x, gridDim.x);
        mf = pol.mf;
    } else {
        const uint32_t *fr = a.fa.bm[L % 3];
        uint32_t *nx = a.fa.bm[(L + 1) % 3];
        int *qout = a.fa.q[(L + 1) % 3];
        int *qsize = &c->qsize[(L + 1) % 3];
        const int lane = lane_id();
        const int wave = (blockIdx.x * BLK + threadIdx.x) / WAVE, nwaves = gridDim.x * BLK / WAVE;
        const int n = a.fa.n;
        for (int base = wave * WAVE; base < n; base += nwaves * WAVE) {
            const int v = base + lane;
            const int w = (base >> 5) + (lane >> 5);
            const uint32_t vw = a.visited[w];
            const bool cand = v < n && !((vw >> (lane & 31)) & 1u);
            bool found = false, hard = false;
            if (cand) {
                const int beg = a.in_off[v], end = a.in_off[v + 1];
                const int lim = min(end, beg + BU_INLINE);
                for (int e = beg; e < lim; ++e)
                    if (test_bit(fr, a.in_adj[e])) {
                        found = true;
                        break;
                    }
                hard = !found && end > lim;
            }
            if (found) {
                a.depth[v] = L + 1;
                mf += a.row_off[v + 1] - a.row_off[v];
            }
            const uint64_t fm = wave_ballot(found);
            const uint32_t mine = (lane < 32) ? (uint32_t)fm : (uint32_t)(fm >> 32);
            if ((lane & 31) == 0 && mine) {  // this wave owns words w of its 64 vertices
                nx[w] = mine;
                a.visited[w] = vw | mine;
            }
            wave_append(found, v, qout, qsize);
            wave_append(hard, v, a.hard, &c->nhard);
        }
    }
    mf = wave_sum_ll(mf);
    if (lane_id() == 0 && mf) atomicAdd((unsigned long long *)&c->mf_next, (unsigned long long)mf);
}

// Edge chunks (top-down) or long in-edge scans (bottom-up) of the level; the last
// workgroup to finish closes the level.
__global__ __launch_bounds__(BLK) void bfs_level_b(BfsArgs a) {
    Ctl *c = a.fa.ctl;
    if (c->done) return;
    const int L = c->level;
    long long mf = 0;
    if (c->dir == DIR_TD) {
        BfsPolicy pol(a);
        td_chunks<BLK>(a.fa, pol, L);
        mf = pol.mf;
    } else {
        const uint32_t *fr = a.fa.bm[L % 3];
        uint32_t *nx = a.fa.bm[(L + 1) % 3];
        const int lane = lane_id();
        const int wave = (blockIdx.x * BLK + threadIdx.x) / WAVE, nwaves = gridDim.x * BLK / WAVE;
        const int nh = c->nhard;
        for (int h = wave; h < nh; h += nwaves) {
            const int v = a.hard[h];
            const int end = a.in_off[v + 1];
            bool found = false;
            for (int e0 = a.in_off[v] + BU_INLINE; e0 < end && !found; e0 += WAVE) {
                const int e = e0 + lane;
                const bool hit = e < end && test_bit(fr, a.in_adj[e]);
                found = wave_ballot(hit) != 0;
            }
            if (found && lane == 0) {
                const uint32_t bit = 1u << (v & 31);
                a.depth[v] = L + 1;
                atomicOr(&a.visited[v >> 5], bit);
                atomicOr(&nx[v >> 5], bit);
                a.fa.q[(L + 1) % 3][atomicAdd(&c->qsize[(L + 1) % 3], 1)] = v;
                mf += a.row_off[v + 1] - a.row_off[v];
            }
        }
    }
    mf = wave_sum_ll(mf);
    if (lane_id() == 0 && mf) atomicAdd((unsigned long long *)&c->mf_next, (unsigned long long)mf);
    if (last_block(c)) {
        bfs_level_end(a);
        __threadfence();
    }
}

}  // namespace mg

using namespace mg;

extern "C" int mg_bfs(mg_graph *g, int32_t src, int32_t *depth_out, const mg_options *opt, mg_stats *stats) {
    CallTimer timer;
    mg_options o;
    int rc = begin_search(g, opt, &o);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g->mu);
    if (src < 0 || src >= g->n) return set_error(MG_ERR_INVALID_ARG, "mg_bfs: source %d outside [0,%d)", src, g->n);
    Workspace &ws = g->ws;
    BfsArgs a{};
    a.fa = make_frontier_args(g, o.directed);
    // Bottom-up needs the in-edges of the traversed graph: the CSR itself when it is
    // symmetric, the CSC for a directed search.
    a.in_off = nullptr;
    a.in_adj = nullptr;
    if (g->flags & MG_GRAPH_SYMMETRIC) {
        a.in_off = g->row_off;
        a.in_adj = g->col;
    } else if (o.directed == 1 && g->col_off) {
        a.in_off = g->col_off;
        a.in_adj = g->row_idx;
    }
    const int can_bu = a.in_off != nullptr && o.engine != MG_ENGINE_GAS;
    a.row_off = g->row_off;
    a.depth = ws.vals;
    a.visited = ws.visited;
    a.hard = ws.hard;
    const long long m_views = g->m * a.fa.nviews;
    const int G = grid_blocks(std::max<int64_t>(g->n / 4, g->m / 64), BLK);

    MG_HIP(hipEventRecord(g->ev0, g->stream));
    hipLaunchKernelGGL(bfs_init, dim3(grid_blocks(g->n, BLK)), dim3(BLK), 0, g->stream, a);
    hipLaunchKernelGGL(bfs_seed, dim3(1), dim3(1), 0, g->stream, a, src, o.iter_num, can_bu, o.alpha, o.beta,
                       m_views);
    MG_HIP(hipGetLastError());
    auto step = [&]() -> int {
        hipLaunchKernelGGL(bfs_small, dim3(1), dim3(SMALL_BLK), 0, g->stream, a);
        for (int v = 0; v < a.fa.nviews; ++v)
            hipLaunchKernelGGL(bfs_level_a, dim3(G), dim3(BLK), 0, g->stream, a, v);
        hipLaunchKernelGGL(bfs_level_b, dim3(G), dim3(BLK), 0, g->stream, a);
        MG_HIP(hipGetLastError());
        return MG_OK;
    };
    const int max_steps = (int)std::min<int64_t>((int64_t)o.iter_num, (int64_t)g->n) + 2;
    if ((rc = drive(g, step, max_steps))) return rc;
    return finish_search(g, stats, ws.vals, -1, depth_out, ws.vals, 4 * (size_t)g->n, o.out_on_device, timer);
}
