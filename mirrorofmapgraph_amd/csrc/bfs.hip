// Breadth-first search: the engine's specialised path for the bfs program
// (Algorithms/BFS/bfs.h; enactor loop GASengine/enactor_vertex_centric.cuh:2526-2872).
//
// The reference runs expand (expand_atomic, LDS/warp/CTA gathering) then contract
// (contract_atomic: label check-and-set, bitmask/hash culls) per level with a host
// round trip per level.  Here the level loop lives on the device; one step is
//   * bfs_small    one 1024-thread workgroup: closes the previous multi-workgroup
//                  level, then runs whole top-down levels back to back while the
//                  frontier is small (frontier queues in HBM, counters in LDS);
//                  long-diameter graphs spend almost all levels here;
//   * bfs_level_a  one level over the whole GPU: top-down (frontier bitmap slices
//                  compacted into LDS, edges gathered through an LDS prefix scan,
//                  hubs split into edge chunks) or bottom-up (every unvisited
//                  vertex scans its in-edges for a parent in the frontier bitmap,
//                  which stays L2-resident: 512 KB at 4M vertices);
//   * bfs_level_b  the edge chunks / long bottom-up scans of the level.
// The direction follows Beamer, Asanovic and Patterson (SC'12): bottom-up once the
// next frontier's out-edges exceed the unexplored edges / alpha, top-down again
// when the frontier falls below n / beta.  Depths are unique, so any traversal
// order yields the reference's labels (bfs.h:241-253: label = iteration + 1 on
// first visit; bfs.cu:56-143 on the CPU).
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
    int G;               // workgroups of the level kernels
};

// Visit of edge (v, u) at level L: claim u through the visited bitmap, label it and
// mark it in the next-frontier bitmap.
struct BfsPolicy {
    const BfsArgs &a;
    long long nf = 0, mf = 0;
    __device__ explicit BfsPolicy(const BfsArgs &args) : a(args) {}
    __device__ int load(int) const { return 0; }
    __device__ bool claim(int u, int L) {
        if (!set_bit_first(a.visited, u)) return false;
        a.depth[u] = L + 1;
        atomicOr(&a.fa.bm[(L + 1) % 3][u >> 5], 1u << (u & 31));
        ++nf;
        mf += view_degree(a.fa, u);
        return true;
    }
    __device__ bool visit(int, int, int, int, int u, bool valid, int L) {
        return valid && !test_bit(a.visited, u) && claim(u, L);
    }
    // Four neighbour ids, then their visited words, then the claims (test-and-set of
    // the visited bits, claim_batch) with the degrees of the claimed vertices.
    __device__ void visit_range(int view, int, int, int beg, int end, int stride, int L) {
        const FrontierArgs &fa = a.fa;
        const int *adj = fa.adj[view];
        uint32_t *nx = fa.bm[(L + 1) % 3];
        for (int e = beg; e < end; e += 4 * stride) {
            int u[4], uu[4];
            load4(adj, e, end, stride, u);
            uint32_t w[4], need = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uu[k] = u[k] >= 0 ? u[k] : 0;  // in bounds; its result is ignored
                w[k] = a.visited[uu[k] >> 5];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (u[k] >= 0 && !((w[k] >> (u[k] & 31)) & 1u)) need |= 1u << k;
            claim_batch<4, true>(fa, a.visited, uu, need, nf, mf, [&](int x) { a.depth[x] = L + 1; }, nx);
        }
    }
};

// Closes level L given the size and out-edges of frontier L+1 (one thread).
__device__ void bfs_close(const BfsArgs &a, int L, long long nf, long long mf) {
    Ctl *c = a.fa.ctl;
    c->m_unexplored -= mf;
    int dir = c->dir;
    trace_level(a.fa, L, dir, nf, mf);
    if (dir == DIR_BU) c->bu_levels++;
    if (c->can_bu) {
        if (dir == DIR_TD && mf > c->m_unexplored / c->alpha) dir = DIR_BU;
        else if (dir == DIR_BU && nf * c->beta < c->n) dir = DIR_TD;
    }
    c->dir = dir;
    advance_level(c, L, nf, mf);
}

// Starts a search in one launch: depths (0 at src, -1 elsewhere), the visited and
// frontier bitmaps (src marked in visited and frontier 0), the level counters, and the
// control block (workgroup 0).  Every word is written once, so no ordering between
// workgroups is needed.
__global__ void bfs_begin(BfsArgs a, int src, int iter_limit, int can_bu, int alpha, int beta, long long m_views) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
    for (int i = tid; i < a.fa.n; i += nt) a.depth[i] = i == src ? 0 : -1;
    const uint32_t bit = 1u << (src & 31);
    for (int i = tid; i < a.fa.nwords; i += nt) {
        const uint32_t b = i == (src >> 5) ? bit : 0u;
        a.visited[i] = b;
        a.fa.bm[0][i] = b;
        a.fa.bm[1][i] = 0;
        a.fa.bm[2][i] = 0;
    }
    if (blockIdx.x != 0) return;
    clear_parts(a.fa);
    if (threadIdx.x != 0) return;
    Ctl *c = a.fa.ctl;
    a.fa.q[0][0] = src;
    ctl_reset(c, a.fa.n, iter_limit, 1);
    const long long d = view_degree(a.fa, src);
    c->can_bu = can_bu;
    c->alpha = alpha > 0 ? alpha : 14;
    c->beta = beta > 0 ? beta : 24;
    c->m_unexplored = m_views - d;
    c->mf_cur = d;
}

// Closes a pending multi-workgroup level, then runs small top-down levels.
//
// Narrow levels are latency-bound (a grid has thousands of them), so the loop keeps
// the chain of dependent memory accesses per level short: the level state lives in
// LDS (the control block is written only when the loop exits), the frontier queue
// lives in LDS, each thread keeps its vertices of the next level and their edge
// offsets in registers (loading them also gives the next frontier's out-edges for
// the direction test), visits claim the visited bit with one atomic (no pre-test;
// workgroup scope: only this workgroup touches the bitmaps during the launch), and a
// frontier's bitmap words are cleared as its vertices are expanded.  Vertices with
// more than LANE_T edges go through the LDS-scan tile path.

__global__ __launch_bounds__(SMALL_BLK) void bfs_small(BfsArgs a) {
    Ctl *c = a.fa.ctl;
    if (threadIdx.x == 0 && !c->done) c->steps++;  // read by the host to size the next search's launches
    const FrontierArgs &fa = a.fa;
    __shared__ int s_q[SMALL_Q];
    __shared__ int s_big[SMALL_BIG];
    __shared__ int s_bpre[SMALL_BIG], s_bbeg[SMALL_BIG], s_bwave[SMALL_BLK / WAVE];
    __shared__ int s_go, s_L, s_cnt, s_nbig, s_levels, s_dir, s_canbu, s_alpha, s_limit;
    __shared__ long long s_mu;
    if (c->pending) {
        const int L = c->level;
        long long nf, mf, ex;
        take_level_parts<SMALL_BLK>(fa, nf, mf, ex);
        if (threadIdx.x == 0) {
            c->edges_examined += ex;
            bfs_close(a, L, nf, mf);
            c->pending = 0;
            c->qvalid = 0;
        }
        __syncthreads();
    }
    int n0 = 0, qv = 0;
    if (threadIdx.x == 0) {
        const int L = c->level;
        s_go = !c->done && c->dir == DIR_TD && c->qsize[L % 3] <= SMALL_Q && c->mf_cur <= SMALL_E;
        s_L = L;
        s_mu = c->m_unexplored;
        s_dir = c->dir;
        s_cnt = c->qsize[L % 3];
        s_levels = 0;
        s_canbu = c->can_bu;
        s_alpha = c->alpha;
        s_limit = c->iter_limit;
        s_q[0] = c->qvalid;  // scratch: queue-valid bits
    }
    __syncthreads();
    if (!s_go) return;
    {
        const int L = s_L, cur = L % 3, prv = (L + 2) % 3;
        qv = s_q[0];
        n0 = s_cnt;
        __syncthreads();
        // this level's frontier into the LDS queue
        if (qv & (1 << cur)) {
            for (int i = threadIdx.x; i < n0; i += SMALL_BLK) s_q[i] = fa.q[cur][i];
        } else {
            if (threadIdx.x == 0) s_cnt = 0;
            __syncthreads();
            queue_from_bitmap_to<SMALL_BLK>(fa, L, s_q, &s_cnt, SMALL_Q);
            n0 = min(s_cnt, SMALL_Q);  // equals qsize[cur] <= SMALL_Q
        }
        // the previous frontier's bitmap is cleared once here; later levels clear
        // their own bitmap as they expand it
        if (qv & (1 << prv)) {
            const int *qp = fa.q[prv];
            for (int i = threadIdx.x; i < c->qsize[prv]; i += SMALL_BLK) fa.bm[prv][qp[i] >> 5] = 0;
        } else {
            clear_prev_sweep(fa, L, threadIdx.x, SMALL_BLK);
        }
        __syncthreads();
    }
    // registers: this thread's frontier vertices and, per view, their edge ranges
    int my_v[SMALL_PER], my_beg[2][SMALL_PER], my_deg[2][SMALL_PER];
    int nmine = 0;
    auto load_mine = [&](int count) -> long long {
        long long mf = 0;
        nmine = 0;
#pragma unroll
        for (int k = 0; k < SMALL_PER; ++k) {
            const int i = (int)threadIdx.x + k * SMALL_BLK;
            my_v[k] = i < count ? s_q[i] : -1;
        }
#pragma unroll
        for (int k = 0; k < SMALL_PER; ++k) {
            for (int view = 0; view < 2; ++view) {
                my_beg[view][k] = 0;
                my_deg[view][k] = 0;
                if (my_v[k] >= 0 && view < fa.nviews) {
                    const int *off = fa.off[view];
                    my_beg[view][k] = off[my_v[k]];
                    my_deg[view][k] = off[my_v[k] + 1] - my_beg[view][k];
                    mf += my_deg[view][k];
                }
            }
            nmine += my_v[k] >= 0;
        }
        return mf;
    };
    load_mine(n0);
    __syncthreads();  // every entry is in registers: the queue is free for the next level
    long long mf_over = 0;  // out-edges of claimed vertices that did not fit the queue
    // Edge slots sl[j] (where ok[j]; sl[j] is in bounds either way): the neighbour ids,
    // then the claims (workgroup-scope test-and-set: this workgroup owns the bitmaps
    // while the loop runs), each stage in flight together; the claimed vertices are
    // labelled, marked and queued.
    auto claim4 = [&](const int *adj, const int (&sl)[4], const bool (&ok)[4], int L, uint32_t *nx) {
        int u[4];
        uint32_t old[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) u[j] = adj[sl[j]];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (ok[j])
                old[j] = __hip_atomic_fetch_or(&a.visited[u[j] >> 5], 1u << (u[j] & 31), __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
        bool over[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            over[j] = false;
            if (ok[j] && !((old[j] >> (u[j] & 31)) & 1u)) {
                a.depth[u[j]] = L + 1;
                __hip_atomic_fetch_or(&nx[u[j] >> 5], 1u << (u[j] & 31), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_WORKGROUP);
                const int idx = atomicAdd(&s_cnt, 1);
                if (idx < SMALL_Q) s_q[idx] = u[j];
                else over[j] = true;
            }
        }
        // claimed vertices past the queue's capacity: their degrees, loaded together
        // (vertex 0's where not needed)
        if (!wave_ballot(over[0] || over[1] || over[2] || over[3])) return;
        int o0[4], o1[4];
#pragma unroll
        for (int view = 0; view < 2; ++view) {
            if (view >= fa.nviews) break;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int d = over[j] ? u[j] : 0;
                o0[j] = fa.off[view][d];
                o1[j] = fa.off[view][d + 1];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (over[j]) mf_over += o1[j] - o0[j];
        }
    };
    for (;;) {
        const int L = s_L;
        uint32_t *nx = fa.bm[(L + 1) % 3];
        uint32_t *cb = fa.bm[L % 3];
        if (threadIdx.x == 0) {
            s_cnt = 0;
            s_nbig = 0;
        }
        __syncthreads();
        // lane path: a thread walks its vertices' lists, four claims in flight
        for (int k = 0; k < SMALL_PER; ++k) {
            const int v = my_v[k];
            if (v < 0) continue;
            cb[v >> 5] = 0;
            if (my_deg[0][k] > LANE_T || my_deg[1][k] > LANE_T) {
                const int bi = atomicAdd(&s_nbig, 1);  // < SMALL_BIG: the frontier has <= SMALL_E edges
                if (bi < SMALL_BIG) s_big[bi] = v;
                else atomicOr(&c->changed, 1 << 30);
                continue;
            }
            for (int view = 0; view < fa.nviews; ++view) {
                const int *adj = fa.adj[view];
                const int b = my_beg[view][k], e = b + my_deg[view][k];
                for (int p = b; p < e; p += 4) {
                    int u[4];
                    uint32_t old[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) u[j] = adj[p + j < e ? p + j : p];
#pragma unroll
                    for (int j = 0; j < 4; ++j) u[j] = p + j < e ? u[j] : -1;
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (u[j] >= 0)
                            old[j] = __hip_atomic_fetch_or(&a.visited[u[j] >> 5], 1u << (u[j] & 31),
                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const bool newly = u[j] >= 0 && !((old[j] >> (u[j] & 31)) & 1u);
                        if (newly) {
                            a.depth[u[j]] = L + 1;
                            __hip_atomic_fetch_or(&nx[u[j] >> 5], 1u << (u[j] & 31), __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_WORKGROUP);
                            const int idx = atomicAdd(&s_cnt, 1);
                            if (idx < SMALL_Q) s_q[idx] = u[j];
                            else mf_over += view_degree(fa, u[j]);
                        }
                    }
                }
            }
        }
        __syncthreads();
        // the few long lists, flattened: the workgroup walks their concatenation, four
        // slots per thread in flight, each slot's list found by binary search over the
        // prefix sums of the lengths
        const int nbig = min(s_nbig, SMALL_BIG);
        for (int view = 0; view < fa.nviews && nbig > 0; ++view) {
            const int *off = fa.off[view];
            int len = 0, beg = 0;
            if ((int)threadIdx.x < nbig) {
                const int v = s_big[threadIdx.x];
                beg = off[v];
                len = off[v + 1] - beg;
            }
            int total;
            const int ex = block_exclusive_scan<SMALL_BLK>(len, s_bwave, &total);
            if ((int)threadIdx.x < nbig) {
                s_bpre[threadIdx.x] = ex;
                s_bbeg[threadIdx.x] = beg;
            }
            __syncthreads();
            for (int p0 = 0; p0 < total; p0 += 4 * SMALL_BLK) {
                int sl[4];
                bool ok[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    int p = p0 + j * SMALL_BLK + (int)threadIdx.x;
                    ok[j] = p < total;
                    p = ok[j] ? p : total - 1;  // past the end: a real slot, not claimed
                    int o = 0;  // last list whose prefix is <= p (a non-empty one)
                    for (int st = 1 << (31 - __builtin_clz(nbig)); st > 0; st >>= 1)
                        if (o + st < nbig && s_bpre[o + st] <= p) o += st;
                    sl[j] = s_bbeg[o] + (p - s_bpre[o]);
                }
                claim4(fa.adj[view], sl, ok, L, nx);
            }
            __syncthreads();  // s_bpre / s_bbeg are rewritten for the next view
        }
        __syncthreads();
        const int nf = s_cnt;
        long long mf = load_mine(min(nf, SMALL_Q)) + mf_over;
        mf_over = 0;
        mf = block_sum<SMALL_BLK>(mf);
        if (threadIdx.x == 0) {
            // close level L on the LDS copy of the direction state (as bfs_close)
            long long mu = s_mu - mf;
            int dir = s_dir;
            trace_level(fa, L, dir, nf, mf);
            if (s_canbu && mf > mu / s_alpha) dir = DIR_BU;
            s_mu = mu;
            s_dir = dir;
            s_levels++;
            const bool go = nf > 0 && nf <= SMALL_Q && dir == DIR_TD && mf <= SMALL_E && L + 1 < s_limit;
            s_go = go;
            s_L = L + 1;
            if (!go) {
                c->m_unexplored = mu;
                c->dir = dir;
                c->small_levels += s_levels;
                advance_level(c, L, nf, mf);
                c->qvalid = 0;  // the queue stayed in LDS
            }
        }
        __syncthreads();
        if (!s_go) break;
    }
}

// One level over the whole GPU: top-down for `view`, or the bottom-up sweep.
__global__ __launch_bounds__(BLK) void bfs_level_a(BfsArgs a, int view) {
    Ctl *c = a.fa.ctl;
    const FrontierArgs &fa = a.fa;
    if (c->done) return;
    const int L = c->level;
    const int dir = c->dir;
    if (dir == DIR_BU && view != 0) return;
    if (view == 0) clear_prev_sweep(fa, L, blockIdx.x * BLK + threadIdx.x, gridDim.x * BLK);
    BfsPolicy pol(a);
    if (dir == DIR_TD) {
        expand_lanes<BLK>(fa, pol, view, L, fa.bm[L % 3]);
    } else {
        __shared__ int s_wave[BLK / WAVE];
        __shared__ int s_hbase;
        const uint32_t *fr = fa.bm[L % 3];
        uint32_t *nx = fa.bm[(L + 1) % 3];
        const int lane = lane_id();
        const int n = fa.n;
        for (int base = blockIdx.x * BLK; base < n; base += gridDim.x * BLK) {
            const int v = base + (int)threadIdx.x;
            const int w = v >> 5;  // this wave owns the two words of its 64 vertices
            const uint32_t vw = v < n ? a.visited[w] : ~0u;
            const bool cand = v < n && !((vw >> (lane & 31)) & 1u);
            bool found = false, hard = false;
            if (cand) {
                const int beg = a.in_off[v], end = a.in_off[v + 1];
                const int lim = min(end, beg + BU_INLINE);
                // one probe at a time: on R-MAT most vertices find a parent at the
                // first in-neighbour (the lists are sorted, hubs have low ids)
                for (int e = beg; e < lim; ++e)
                    if (test_bit(fr, a.in_adj[e])) {
                        found = true;
                        break;
                    }
                hard = !found && end > lim;
            }
            if (found) {
                a.depth[v] = L + 1;
                ++pol.nf;
                pol.mf += a.row_off[v + 1] - a.row_off[v];
            }
            const uint64_t fm = wave_ballot(found);
            const uint32_t mine = (lane < 32) ? (uint32_t)fm : (uint32_t)(fm >> 32);
            if ((lane & 31) == 0 && mine) {
                nx[w] = mine;
                a.visited[w] = vw | mine;
            }
            int htot;
            const int hex = block_exclusive_scan<BLK>(hard ? 1 : 0, s_wave, &htot);
            if (htot > 0) {
                if (threadIdx.x == 0) s_hbase = atomicAdd(&c->nhard, htot);
                __syncthreads();
                if (hard) a.hard[s_hbase + hex] = v;
                __syncthreads();
            }
        }
    }
    write_part<BLK>(fa, pol.nf, pol.mf, 0);
}

// Edge chunks (top-down) or long in-edge scans (bottom-up) of the level.
__global__ __launch_bounds__(BLK) void bfs_level_b(BfsArgs a) {
    Ctl *c = a.fa.ctl;
    const FrontierArgs &fa = a.fa;
    if (c->done) return;
    const int L = c->level;
    BfsPolicy pol(a);
    if (c->dir == DIR_TD) {
        expand_tasks<BLK>(fa, pol, L);
    } else {
        const uint32_t *fr = fa.bm[L % 3];
        uint32_t *nx = fa.bm[(L + 1) % 3];
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
                ++pol.nf;
                pol.mf += a.row_off[v + 1] - a.row_off[v];
            }
        }
    }
    write_part<BLK>(fa, pol.nf, pol.mf, 0);
    if (blockIdx.x == 0 && threadIdx.x == 0) c->pending = 1;
}

// Predecessors (Algorithms/BFS/predextract.cuh:246-375, run by bfs.cu:459-468): the
// reference gathers, over every vertex's in-edges, the neighbour id when
// label[v] - label[u] == 1 and -1 otherwise, and reduces them with max.  Lists up to
// LANE_T in-edges: one lane each; longer lists: TASK-edge pieces, one wave each,
// folded in with atomicMax.
__global__ __launch_bounds__(BLK) void bfs_pred_lanes(const int *in_off, const int *in_adj, const int *label,
                                                     int *pred, int n) {
    for (int v = blockIdx.x * BLK + (int)threadIdx.x; v < n; v += gridDim.x * BLK) {
        const int beg = in_off[v], end = in_off[v + 1];
        int best = -1;
        if (end - beg <= LANE_T) {
            const int lv = label[v];
            for (int e = beg; e < end; ++e) {
                const int u = in_adj[e];
                if (lv - label[u] == 1) best = max(best, u);
            }
        }
        pred[v] = best;
    }
}

__global__ __launch_bounds__(BLK) void bfs_pred_tasks(const ChunkRec *tasks, int ntasks, const int *in_adj,
                                                     const int *label, int *pred) {
    const int lane = lane_id();
    const int wave = (blockIdx.x * BLK + (int)threadIdx.x) / WAVE, nwaves = gridDim.x * BLK / WAVE;
    for (int t = wave; t < ntasks; t += nwaves) {
        const ChunkRec r = tasks[t];
        const int lv = label[r.v];
        int best = -1;
        for (int e = r.beg + lane; e < r.end; e += WAVE) {
            const int u = in_adj[e];
            if (lv - label[u] == 1) best = max(best, u);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) best = max(best, __shfl_xor(best, o));
        if (lane == 0 && best >= 0) atomicMax(&pred[r.v], best);
    }
}

}  // namespace mg

using namespace mg;

extern "C" int mg_bfs_pred(mg_graph *g, int32_t src, int32_t *depth_out, int32_t *pred_out, const mg_options *opt,
                           mg_stats *stats) {
    CallTimer timer;
    mg_options o;
    int rc = begin_search(g, opt, &o);
    if (rc) return rc;
    const bool sym = g->flags & MG_GRAPH_SYMMETRIC;
    if (!sym && !g->col_off)
        return set_error(MG_ERR_UNSUPPORTED, "mg_bfs_pred: needs the in-edge view (MG_GRAPH_BUILD_CSC or MG_GRAPH_SYMMETRIC)");
    mg_options inner = o;
    inner.out_on_device = 1;
    inner.engine = MG_ENGINE_AUTO;  // the depths must land in this graph's workspace
    inner.layout = 1;               // (not a degree-ordered twin's)
    if ((rc = mg_bfs(g, src, nullptr, &inner, stats))) return rc;  // depths stay in the workspace
    std::lock_guard<std::mutex> lock(g->mu);
    Workspace &ws = g->ws;
    const int *in_off = sym ? g->row_off : g->col_off;
    const int *in_adj = sym ? g->col : g->row_idx;
    if ((rc = ensure_pull_tasks(g, in_off))) return rc;
    int *pred = nullptr;
    if ((rc = alloc_i32(&pred, (size_t)g->n))) return rc;
    hipLaunchKernelGGL(bfs_pred_lanes, dim3(grid_blocks(g->n, BLK)), dim3(BLK), 0, g->stream, in_off, in_adj,
                       (const int *)ws.vals, pred, g->n);
    if (ws.pull_ntasks)
        hipLaunchKernelGGL(bfs_pred_tasks, dim3(grid_blocks((int64_t)ws.pull_ntasks * WAVE, BLK)), dim3(BLK), 0,
                           g->stream, (const ChunkRec *)ws.pull_tasks, ws.pull_ntasks, in_adj, (const int *)ws.vals,
                           pred);
    rc = MG_OK;
    if (hipGetLastError() != hipSuccess) rc = set_error(MG_ERR_DEVICE, "predecessor kernels failed to launch");
    if (!rc) rc = deliver(g, depth_out, ws.vals, 4 * (size_t)g->n, o.out_on_device);
    if (!rc) rc = deliver(g, pred_out, pred, 4 * (size_t)g->n, o.out_on_device);
    (void)hipStreamSynchronize(g->stream);
    free_dev(pred);
    if (stats) stats->host_ms = timer.ms();
    return rc;
}

namespace mg {

// One BFS search with the current queues (MG_ERR_OVERFLOW when the edge-chunk queue
// overflowed; the caller regrows it and runs again).
static int bfs_run(mg_graph *g, int src, int *depth_out, const mg_options &o, mg_stats *stats,
                   const CallTimer &timer) {
    int rc;
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
    a.G = o.level_grid > 0 ? o.level_grid : grid_blocks(std::max<int64_t>(g->n / 4, g->m / 64), BLK);
    const long long m_views = g->m * a.fa.nviews;

    MG_HIP(hipEventRecord(g->ev0, g->stream));
    hipLaunchKernelGGL(bfs_begin, dim3(grid_blocks(g->n, BLK)), dim3(BLK), 0, g->stream, a, src, o.iter_num, can_bu,
                       o.alpha, o.beta, m_views);
    MG_HIP(hipGetLastError());
    auto step = [&]() -> int {
        hipLaunchKernelGGL(bfs_small, dim3(1), dim3(SMALL_BLK), 0, g->stream, a);
        for (int v = 0; v < a.fa.nviews; ++v)
            hipLaunchKernelGGL(bfs_level_a, dim3(a.G), dim3(BLK), 0, g->stream, a, v);
        hipLaunchKernelGGL(bfs_level_b, dim3(a.G), dim3(BLK), 0, g->stream, a);
        MG_HIP(hipGetLastError());
        return MG_OK;
    };
    const int64_t max_steps = std::min<int64_t>((int64_t)o.iter_num, (int64_t)g->n) + 3;
    bool tail_queued = false;
    auto tail = [&]() { return queue_finish(g, stats != nullptr, ws.vals, -1, ws.vals, 4 * (size_t)g->n); };
    if ((rc = drive(g, step, max_steps, KIND_BFS, mix64((uint64_t)o.directed + 2 * (uint64_t)a.G), nullptr, nullptr,
                    tail, &tail_queued)))
        return rc;
    rc = finish_search(g, stats, ws.vals, -1, depth_out, ws.vals, 4 * (size_t)g->n, o.out_on_device, timer,
                       tail_queued);
    if (rc == MG_OK && o.verbose) rc = print_trace(g, "bfs", "top-down", "bottom-up");
    return rc;
}

}  // namespace mg

extern "C" int mg_bfs(mg_graph *g, int32_t src, int32_t *depth_out, const mg_options *opt, mg_stats *stats) {
    CallTimer timer;
    mg_options o;
    int rc = begin_search(g, opt, &o);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g->mu);
    if (src < 0 || src >= g->n) return set_error(MG_ERR_INVALID_ARG, "mg_bfs: source %d outside [0,%d)", src, g->n);
    if (o.engine == MG_ENGINE_GAS) return gas_bfs(g, src, depth_out, o, stats, timer);
    g = search_graph(g, o, &src);  // the degree-ordered twin, when the graph keeps one
    if ((rc = ensure_workspace(g))) return rc;
    if ((rc = size_chunk_queue(g, o))) return rc;
    int regrows = 0;
    while ((rc = bfs_run(g, src, depth_out, o, stats, timer)) == MG_ERR_OVERFLOW && regrows < MAX_REGROWS) {
        ++regrows;
        if ((rc = grow_chunk_queue(g))) return rc;
        clear_error();
    }
    if (rc == MG_OK && stats) stats->overflow_regrows = regrows;
    return rc;
}
