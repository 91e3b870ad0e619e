// Batched multi-source BFS: up to 256 sources (slots) share every edge read.
//
// The reference runs one search per source, re-initialising between them
// (Algorithms/BFS/bfs.cu:402-417).  Here each vertex carries three 256-bit masks
// (one 32-byte sector each; bit s = slot s): V (visited), F (frontier), N (next).
// One level is
//   * push (sparse levels): every active vertex v (F[v] != 0, listed in the active
//     bitmap A[L%2]) ORs F[v] & ~V[u] into N[u] for each out-neighbour u (atomicOr on
//     the words that gain bits) and marks u in A[(L+1)%2]; lists longer than LANE_T
//     become TASK-edge tasks walked by one wave;
//   * pull (dense levels, needs in-edges): every vertex u not yet reached by every
//     assigned slot ORs the frontier masks of its in-neighbours, stopping once every
//     slot has reached it, and stores the new bits in N[u] (no atomics);
//   * finalize: F[v] = 0 for the old frontier, then for every marked u: new = N[u],
//     V[u] |= new, F[u] = new, N[u] = 0, and row[s][u] = L + 1 - start[s] for each
//     new bit s (start[s] = the level the slot was seeded at).
// Depths are unique per source, so row s equals a single-source search from the
// slot's source, whatever other slots run beside it and whenever it was seeded.
// Level close, direction choice and termination run on the device in the next
// step's single-workgroup kernel; steps are replayed from a captured HIP graph.
// Slots can be seeded between steps (serving): a new slot joins the next level's
// frontier; a slot whose bit left every frontier mask is finished.
#include "batch.h"

namespace mg {

struct BatchArgs {
    Ctl *ctl;
    SlotCtl *sc;
    const int *off, *adj;        // push view (CSR)
    const int *in_off, *in_adj;  // pull view (in-edges), nullptr when none
    uint4 *V, *F, *N;            // 2 uint4 per vertex
    uint32_t *act[2];            // active-vertex bitmaps
    int *out;                    // S x n depths
    int *hard;                   // pull: vertices handed to a wave
    ChunkRec *tasks;
    int64_t task_cap;
    LevelPart *part;
    uint4 *slot_or;              // per-workgroup OR of new frontier masks (2 uint4 each)
    int n, nwords, G;
};

__device__ __forceinline__ M8 zero8() { return M8{make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)}; }

// Block-wide OR of a mask; the result is valid in thread 0.
template <int BLOCK>
__device__ M8 block_or8(M8 x) {
    __shared__ uint4 s_or[2 * (BLOCK / WAVE)];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = or8(x, shfl_xor8(x, o));
    const int wv = (int)threadIdx.x / WAVE;
    if (lane_id() == 0) {
        s_or[2 * wv] = x.a;
        s_or[2 * wv + 1] = x.b;
    }
    __syncthreads();
    M8 r = zero8();
    if (threadIdx.x == 0)
        for (int i = 0; i < BLOCK / WAVE; ++i) r = or8(r, M8{s_or[2 * i], s_or[2 * i + 1]});
    __syncthreads();
    return r;
}

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

__global__ void batch_init(BatchArgs a, int iter_limit, int can_pull, int alpha, int beta, long long m,
                           int rows_to_clear) {
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, nt = (int64_t)gridDim.x * blockDim.x;
    const int64_t kn = (int64_t)rows_to_clear * a.n;
    for (int64_t i = tid; i < kn; i += nt) a.out[i] = -1;
    for (int64_t i = tid; i < 2 * (int64_t)a.n; i += nt) {
        a.V[i] = make_uint4(0, 0, 0, 0);
        a.F[i] = make_uint4(0, 0, 0, 0);
        a.N[i] = make_uint4(0, 0, 0, 0);
    }
    for (int64_t i = tid; i < a.nwords; i += nt) {
        a.act[0][i] = 0;
        a.act[1][i] = 0;
    }
    int *sci = reinterpret_cast<int *>(a.sc);
    for (int64_t i = tid; i < (int64_t)(sizeof(SlotCtl) / 4); i += nt) sci[i] = 0;
    for (int64_t i = tid; i < 2 * MAX_GRID; i += nt) a.slot_or[i] = make_uint4(0, 0, 0, 0);
    if (tid == 0) {
        Ctl *c = a.ctl;
        ctl_reset(c, a.n, iter_limit, 0);  // idle until a slot is seeded
        c->can_bu = can_pull;
        c->m_unexplored = m;  // pull once the frontier's out-edges exceed m / alpha
        c->alpha = alpha;
        c->beta = beta;
    }
}

// Clears the visited bits and rows of the slots about to be seeded.
__global__ void batch_prepare(BatchArgs a, M8 newmask, SlotList sl) {
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, nt = (int64_t)gridDim.x * blockDim.x;
    for (int64_t u = tid; u < a.n; u += nt) {
        const M8 v = ld8(a.V, (int)u);
        const M8 k = andnot8(v, newmask);
        if (any8(andnot8(v, k))) st8(a.V, (int)u, k);
        for (int j = 0; j < sl.n; ++j) a.out[(int64_t)sl.slot[j] * a.n + u] = -1;
    }
}

// One thread per (slot, source) pair.  A slot seeded while a level is pending
// joins the next level's frontier; seeded into an idle context it starts the search.
__global__ void batch_seed(BatchArgs a, SeedList sl) {
    Ctl *c = a.ctl;
    const int i = threadIdx.x;
    const int pending = c->pending;
    const int Lf = pending ? c->level + 1 : c->level;
    if (i < sl.n) {
        const int s = sl.slot[i], v = sl.src[i];
        const uint32_t sb = 1u << (s & 31);
        atomicOr(reinterpret_cast<uint32_t *>(a.V + 2 * v) + (s >> 5), sb);
        atomicOr(reinterpret_cast<uint32_t *>(a.F + 2 * v) + (s >> 5), sb);
        atomicOr(&a.act[Lf & 1][v >> 5], 1u << (v & 31));
        atomicOr(&a.sc->assigned[s >> 5], sb);
        atomicOr(&a.sc->live[s >> 5], sb);
        a.sc->start[s] = Lf;
        a.sc->te[s] = 0;
        a.sc->nr[s] = 0;
        a.out[(int64_t)s * a.n + v] = 0;
        atomicAdd((unsigned long long *)&c->edges_traversed, (unsigned long long)(a.off[v + 1] - a.off[v]));
        atomicAdd((unsigned long long *)&c->reached, 1ULL);
    }
    __syncthreads();
    if (i == 0 && sl.n > 0) {
        if (pending) atomicAdd((unsigned long long *)&a.part[0].nf, (unsigned long long)sl.n);
        else c->done = c->iter_limit <= 0;
    }
}

// One workgroup: closes the pending level (partials of batch_fin).
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
        const M8 full = ld8(reinterpret_cast<const uint4 *>(a.sc->assigned), 0);
        for (int base = blockIdx.x * BLK; base < a.n; base += gridDim.x * BLK) {
            const int u = base + (int)threadIdx.x;
            bool found = false, hard = false;
            if (u < a.n) {
                const M8 vu = ld8(a.V, u);
                if (any8(andnot8(full, vu))) {
                    M8 acc = zero8();
                    const int beg = a.in_off[u], end = a.in_off[u + 1];
                    const int lim = min(end, beg + BU_INLINE);
                    int e = beg;
                    for (; e < lim; ++e) {
                        acc = or8(acc, ld8(a.F, a.in_adj[e]));
                        if (!any8(andnot8(full, or8(acc, vu)))) break;
                    }
                    const M8 nw = andnot8(acc, vu);
                    if (any8(nw)) {
                        st8(a.N, u, nw);
                        found = true;
                    }
                    hard = e == lim && lim < end && any8(andnot8(full, or8(acc, vu)));
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
            M8 acc = zero8();
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

// Finalize: the old frontier's masks are cleared (except for vertices of the new
// frontier, whose masks are overwritten below, so no vertex is written twice) along
// with the old active bitmap, which becomes the next level's empty target; new bits
// become visited and the next frontier; depths are written.
__global__ __launch_bounds__(BLK) void batch_fin(BatchArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const int L = c->level;
    const uint32_t *nxt = a.act[(L + 1) & 1];
    for_each_bit_clear(a.act[L & 1], a.nwords, [&](int v) {
        if (v >= 0 && !test_bit(nxt, v)) st8(a.F, v, zero8());
    });
    long long nf = 0, mf = 0, te = 0, nr = 0;
    M8 lv = zero8();
    for_each_bit(a.act[(L + 1) & 1], a.nwords, [&](int u) {
        if (u < 0) return;
        const M8 nw = ld8(a.N, u);
        st8(a.N, u, zero8());
        st8(a.V, u, or8(ld8(a.V, u), nw));
        st8(a.F, u, nw);
        lv = or8(lv, nw);
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
                a.out[(int64_t)s * a.n + u] = L + 1 - a.sc->start[s];
                b &= b - 1;
            }
        }
    });
    nf = block_sum<BLK>(nf);
    mf = block_sum<BLK>(mf);
    te = block_sum<BLK>(te);
    nr = block_sum<BLK>(nr);
    lv = block_or8<BLK>(lv);
    if (threadIdx.x == 0) {
        LevelPart p;
        p.nf = nf;
        p.mf = mf;
        p.ex = te;
        p.pad = nr;  // (source, vertex) pairs reached
        a.part[blockIdx.x] = p;
        st8(a.slot_or, blockIdx.x, lv);
        if (blockIdx.x == 0) c->pending = 1;
    }
}

// ---------------------------------------------------------------- slot utilities

// One workgroup: ORs the per-workgroup frontier masks written by the last
// finalize into sc->live.  Run right after a run of steps, before any seeding.
__global__ __launch_bounds__(SMALL_BLK) void k_batch_live(Ctl *c, SlotCtl *sc, const uint4 *slot_or, int G) {
    (void)c;
    M8 x = zero8();
    for (int i = threadIdx.x; i < G; i += SMALL_BLK) x = or8(x, ld8(slot_or, i));
    x = block_or8<SMALL_BLK>(x);
    if (threadIdx.x == 0) st8(reinterpret_cast<uint4 *>(sc->live), 0, x);
}

// Grid (ns, B): workgroup (j, b) counts a stripe of row slot[j]: vertices reached
// and their out-degrees, one atomic per workgroup.
__global__ __launch_bounds__(BLK) void k_row_stats(const int *rows, const int *off, int n, SlotList sl, int unreached,
                                                    SlotCtl *sc) {
    const int s = sl.slot[blockIdx.y];
    const int *row = rows + (int64_t)s * n;
    long long te = 0, nr = 0;
    for (int u = blockIdx.x * BLK + (int)threadIdx.x; u < n; u += gridDim.x * BLK)
        if (row[u] != unreached) {
            ++nr;
            te += off[u + 1] - off[u];
        }
    te = block_sum<BLK>(te);
    nr = block_sum<BLK>(nr);
    if (threadIdx.x == 0) {
        atomicAdd((unsigned long long *)&sc->te[s], (unsigned long long)te);
        atomicAdd((unsigned long long *)&sc->nr[s], (unsigned long long)nr);
    }
}

__global__ void k_release_slots(SlotCtl *sc, M8 mask) {
    const int w = threadIdx.x;
    if (w < 8) {
        sc->assigned[w] &= ~word8(mask, w);
        sc->live[w] &= ~word8(mask, w);
    }
}

static M8 host_mask(const uint32_t m[8]) {
    return M8{make_uint4(m[0], m[1], m[2], m[3]), make_uint4(m[4], m[5], m[6], m[7])};
}

static BatchArgs bfs_args(BatchDev &b, mg_graph *g) {
    BatchArgs a{};
    a.ctl = b.ctl;
    a.sc = b.sc;
    a.off = g->row_off;
    a.adj = g->col;
    if (g->flags & MG_GRAPH_SYMMETRIC) {
        a.in_off = g->row_off;
        a.in_adj = g->col;
    } else if (g->col_off) {
        a.in_off = g->col_off;
        a.in_adj = g->row_idx;
    }
    a.V = b.V;
    a.F = b.F;
    a.N = b.N;
    a.act[0] = b.act[0];
    a.act[1] = b.act[1];
    a.out = b.rows;
    a.hard = b.hard;
    a.tasks = b.tasks;
    a.task_cap = b.task_cap;
    a.part = b.part;
    a.slot_or = b.slot_or;
    a.n = g->n;
    a.nwords = b.nwords;
    a.G = b.G;
    return a;
}

int bfs_batch_reset(BatchDev &b, mg_graph *g, hipStream_t st, int iter_limit, int alpha, int beta,
                    int rows_to_clear) {
    BatchArgs a = bfs_args(b, g);
    const int can_pull = a.in_off != nullptr;
    hipLaunchKernelGGL(batch_init, dim3(grid_blocks(std::max<int64_t>((int64_t)rows_to_clear * g->n, 2 * g->n), BLK)),
                       dim3(BLK), 0, st, a, iter_limit, can_pull, alpha, beta, (long long)g->m, rows_to_clear);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int bfs_batch_seed(BatchDev &b, mg_graph *g, hipStream_t st, const SeedList &sl) {
    if (sl.n <= 0) return MG_OK;
    BatchArgs a = bfs_args(b, g);
    uint32_t m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    SlotList ls{};
    ls.n = sl.n;
    for (int i = 0; i < sl.n; ++i) {
        m[sl.slot[i] >> 5] |= 1u << (sl.slot[i] & 31);
        ls.slot[i] = sl.slot[i];
    }
    hipLaunchKernelGGL(batch_prepare, dim3(grid_blocks(g->n, BLK)), dim3(BLK), 0, st, a, host_mask(m), ls);
    hipLaunchKernelGGL(batch_seed, dim3(1), dim3(MG_BATCH_MAX), 0, st, a, sl);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int bfs_batch_seed_pairs(BatchDev &b, mg_graph *g, hipStream_t st, const int32_t *slot_of, const int32_t *srcs,
                         int npairs) {
    BatchArgs a = bfs_args(b, g);
    uint32_t m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    SlotList ls{};
    for (int i = 0; i < npairs; ++i) {
        const int s = slot_of[i];
        if (!((m[s >> 5] >> (s & 31)) & 1u)) ls.slot[ls.n++] = s;
        m[s >> 5] |= 1u << (s & 31);
    }
    // the slots' old state is cleared once, then every pair sets its source (a slot's
    // sources together form its frontier 0)
    hipLaunchKernelGGL(batch_prepare, dim3(grid_blocks(g->n, BLK)), dim3(BLK), 0, st, a, host_mask(m), ls);
    for (int i0 = 0; i0 < npairs; i0 += MG_BATCH_MAX) {
        SeedList sl{};
        for (int i = i0; i < npairs && i < i0 + MG_BATCH_MAX; ++i) {
            sl.slot[sl.n] = slot_of[i];
            sl.src[sl.n++] = srcs[i];
        }
        hipLaunchKernelGGL(batch_seed, dim3(1), dim3(MG_BATCH_MAX), 0, st, a, sl);
    }
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int bfs_batch_step(BatchDev &b, mg_graph *g, hipStream_t st) {
    BatchArgs a = bfs_args(b, g);
    hipLaunchKernelGGL(batch_close, dim3(1), dim3(SMALL_BLK), 0, st, a);
    hipLaunchKernelGGL(batch_a, dim3(a.G), dim3(BLK), 0, st, a);
    hipLaunchKernelGGL(batch_b, dim3(a.G), dim3(BLK), 0, st, a);
    hipLaunchKernelGGL(batch_fin, dim3(a.G), dim3(BLK), 0, st, a);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int batch_live(BatchDev &b, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_live, dim3(1), dim3(SMALL_BLK), 0, st, b.ctl, b.sc, (const uint4 *)b.slot_or, b.G);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int batch_row_stats(BatchDev &b, mg_graph *g, hipStream_t st, const SlotList &sl, int unreached) {
    if (sl.n <= 0) return MG_OK;
    const int bx = std::max(1, std::min(64, grid_blocks(g->n, BLK)));
    hipLaunchKernelGGL(k_row_stats, dim3(bx, sl.n), dim3(BLK), 0, st, (const int *)b.rows, (const int *)g->row_off,
                       g->n, sl, unreached, b.sc);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int batch_release_slots(BatchDev &b, hipStream_t st, const uint32_t mask[8]) {
    hipLaunchKernelGGL(k_release_slots, dim3(1), dim3(64), 0, st, b.sc, host_mask(mask));
    MG_HIP(hipGetLastError());
    return MG_OK;
}

// ---------------------------------------------------------------- device state

int64_t batch_dev_bytes(const mg_graph *g, bool sssp, int slots) {
    const int64_t n = g->n, nwords = ((n + 63) / 64) * 2;
    const int KS = (slots + 15) / 16 * 16;
    int64_t b = (int64_t)sizeof(Ctl) + sizeof(SlotCtl) + 3LL * MAX_GRID * sizeof(LevelPart) + 32LL * MAX_GRID;
    b += 2 * 4 * nwords + 3 * (32 * n + 32) + 4LL * slots * n + 4;
    if (g->perm) b += 4LL * slots * n + 4;  // rows_out
    if (sssp) {
        b += (32 * n + 32) + 4LL * SB_RING * nwords + 4LL * KS * n + 64 + 8LL * 4 * n;
        b += (int64_t)sizeof(ChunkRec) * (g->m / SB_TASK + n + 64);
    } else {
        b += 4 * n + 4 + (int64_t)sizeof(ChunkRec) * (2 * (g->m / TASK) + 64);
    }
    return b;
}

void batch_dev_free(BatchDev &b) {
    free_dev(b.ctl);
    free_dev(b.sc);
    free_dev(b.part);
    free_dev(b.slot_or);
    free_dev(b.act[0]);
    free_dev(b.act[1]);
    free_dev(b.ring);
    free_dev(b.hard);
    free_dev(b.tasks);
    free_dev(b.V);
    free_dev(b.F);
    free_dev(b.N);
    free_dev(b.FAR);
    free_dev(b.D);
    free_dev(b.pairs);
    free_dev(b.rows);
    free_dev(b.rows_out);
    if (b.h_ctl) (void)hipHostFree(b.h_ctl);
    if (b.h_sc) (void)hipHostFree(b.h_sc);
    b = BatchDev{};
}

int batch_dev_alloc(BatchDev &b, mg_graph *g, bool sssp, int slots) {
    if (slots < 1 || slots > MG_BATCH_MAX) return set_error(MG_ERR_INVALID_ARG, "batch slots must be in [1, %d]", MG_BATCH_MAX);
    batch_dev_free(b);
    const size_t n = (size_t)g->n;
    b.sssp = sssp;
    b.S = slots;
    b.KS = (slots + 15) / 16 * 16;
    b.nwords = (int)(((n + 63) / 64) * 2);
    b.G = sssp ? grid_blocks(std::max<int64_t>(g->n / 4, g->m / 16), BLK) : grid_blocks(std::max<int64_t>(g->n / 4, g->m / 64), BLK);
    b.task_need = sssp ? g->m / SB_TASK + (int64_t)n + 64 : 2 * (g->m / TASK) + 64;
    b.task_cap = b.task_need;
    hipError_t e = hipSuccess;
    auto A = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = alloc_refused() ? hipErrorOutOfMemory : hipMalloc(p, bytes ? bytes : 16);
    };
    A((void **)&b.ctl, sizeof(Ctl));
    A((void **)&b.sc, sizeof(SlotCtl));
    A((void **)&b.part, 3 * MAX_GRID * sizeof(LevelPart));
    A((void **)&b.slot_or, 32 * MAX_GRID);
    A((void **)&b.act[0], 4 * (size_t)b.nwords);
    A((void **)&b.act[1], 4 * (size_t)b.nwords);
    A((void **)&b.tasks, sizeof(ChunkRec) * (size_t)b.task_cap);
    A((void **)&b.F, 32 * n + 32);
    A((void **)&b.N, 32 * n + 32);
    A((void **)&b.rows, 4 * (size_t)slots * n + 4);
    if (g->perm) A((void **)&b.rows_out, 4 * (size_t)slots * n + 4);
    if (sssp) {
        A((void **)&b.FAR, 32 * n + 32);
        A((void **)&b.ring, 4 * (size_t)SB_RING * b.nwords);
        A((void **)&b.D, 4 * (size_t)b.KS * n + 64);
        b.pair_cap = 4 * (int64_t)n;  // more pairs in one iteration fall back to wave tasks
        A((void **)&b.pairs, sizeof(int2) * (size_t)b.pair_cap);
        A((void **)&b.V, 32);  // unused by SSSP
    } else {
        A((void **)&b.V, 32 * n + 32);
        A((void **)&b.hard, 4 * n + 4);
    }
    if (e == hipSuccess) e = hipHostMalloc((void **)&b.h_ctl, sizeof(Ctl), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&b.h_sc, sizeof(SlotCtl), hipHostMallocDefault);
    if (e != hipSuccess) {
        batch_dev_free(b);
        return set_error(MG_ERR_OOM, "batch state allocation (%d slots) failed: %s", slots, hipGetErrorString(e));
    }
    b.bytes = batch_dev_bytes(g, sssp, slots);
    return MG_OK;
}

int64_t batch_dev_size(const BatchDev *b) { return b->bytes; }

static int set_task_cap(mg_graph *g, BatchDev &b, int64_t cap) {
    if (cap == b.task_cap) return MG_OK;
    MG_HIP(hipStreamSynchronize(g->stream));  // no queued step may still use the old queue
    invalidate_steps(g);
    b.bytes -= (int64_t)sizeof(ChunkRec) * b.task_cap;
    free_dev(b.tasks);
    b.tasks = nullptr;
    b.task_cap = 0;
    if (alloc_refused() || hipMalloc((void **)&b.tasks, sizeof(ChunkRec) * (size_t)cap) != hipSuccess) {
        b.tasks = nullptr;
        return set_error(MG_ERR_OOM, "batched task queue of %lld records: allocation failed", (long long)cap);
    }
    b.task_cap = cap;
    b.bytes += (int64_t)sizeof(ChunkRec) * cap;
    return MG_OK;
}

int size_task_queue(mg_graph *g, BatchDev &b, const mg_options &o) {
    return set_task_cap(g, b, std::max(queue_capacity(b.task_need, o.max_queue_sizing), b.task_floor));
}

int grow_task_queue(mg_graph *g, BatchDev &b) {
    b.task_floor = 4 * b.task_cap;
    return set_task_cap(g, b, b.task_floor);
}

void batch_release(mg_graph *g) {
    for (BatchDev *&p : g->steps->bdev) {
        if (p) {
            batch_dev_free(*p);
            delete p;
            p = nullptr;
        }
    }
}

// The graph's own batch context for kind 0 (BFS) or 1 (SSSP), allocated once at
// MG_BATCH_MAX slots so captured step graphs keep valid pointers.
int graph_batch_dev(mg_graph *g, bool sssp, BatchDev **out) {
    BatchDev *&p = g->steps->bdev[sssp ? 1 : 0];
    if (!p) {
        p = new (std::nothrow) BatchDev();
        if (!p) return set_error(MG_ERR_OOM, "host allocation failed");
        int rc = batch_dev_alloc(*p, g, sssp, MG_BATCH_MAX);
        if (rc) {
            delete p;
            p = nullptr;
            return rc;
        }
    }
    *out = p;
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

// out[s * n + v] = rows[s * n + perm[v]]: batched rows of a degree-ordered twin in the
// caller's numbering (grid: x over vertices, y over rows)
__global__ void unpermute_rows(const int *perm, const int *rows, int n, int *out) {
    const int64_t base = (int64_t)blockIdx.y * n;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        out[base + v] = rows[base + perm[v]];
}

int launch_unpermute_rows(const int *perm, const int *rows, int n, int k, int *out, hipStream_t st) {
    if (k <= 0 || n <= 0) return MG_OK;
    hipLaunchKernelGGL(unpermute_rows, dim3(grid_blocks_cap(n, BLK, 64), k), dim3(BLK), 0, st, perm, rows, n, out);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int finish_batch(mg_graph *g, BatchDev &b, mg_stats *stats, void *out, size_t bytes, bool out_on_device,
                 const CallTimer &timer) {
    const int *result = b.rows;
    if (g->perm) {  // a search on a twin: every row back to the caller's ids
        const int k = (int)(bytes / (4 * (size_t)std::max(g->n, 1)));
        if (k > 0 && g->n > 0) {
            hipLaunchKernelGGL(unpermute_rows, dim3(grid_blocks_cap(g->n, BLK, 64), k), dim3(BLK), 0, g->stream,
                               (const int *)g->perm, (const int *)b.rows, g->n, b.rows_out);
            MG_HIP(hipGetLastError());
        }
        result = b.rows_out;
    }
    MG_HIP(hipEventRecord(g->ev1, g->stream));
    MG_HIP(hipEventSynchronize(g->ev1));
    float ms = 0;
    MG_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
    MG_HIP(copy_on(g->stream, b.h_ctl, b.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    if (b.h_ctl->changed & (1 << 30))
        return set_error(MG_ERR_OVERFLOW, "batched task queue overflow (capacity %lld)", (long long)b.task_cap);
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->iterations = b.h_ctl->level;
        stats->bottom_up_levels = b.h_ctl->bu_levels;
        stats->edges_examined = b.h_ctl->edges_examined;
        stats->edges_traversed = b.h_ctl->edges_traversed;
        stats->vertices_reached = b.h_ctl->reached;
        stats->device_ms = ms;
    }
    int rc = deliver(g, out, result, bytes, out_on_device);
    if (rc) return rc;
    if (stats) stats->host_ms = timer.ms();
    return MG_OK;
}

int batch_alpha(const mg_options &o) { return o.batch_alpha > 0 ? o.batch_alpha : 4; }
int batch_beta(const mg_options &o) { return o.batch_beta > 0 ? o.batch_beta : 16; }

}  // namespace mg

using namespace mg;

namespace mg {

int check_batch_sets(const char *fn, const mg_graph *g, const int32_t *slot_of, const int32_t *srcs, int32_t npairs,
                     int32_t k) {
    if (k < 1 || k > MG_BATCH_MAX) return set_error(MG_ERR_INVALID_ARG, "%s: k must be in [1, %d]", fn, MG_BATCH_MAX);
    if (npairs < 1 || !srcs || !slot_of) return set_error(MG_ERR_INVALID_ARG, "%s: no (slot, source) pairs", fn);
    uint32_t seen[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < npairs; ++i) {
        if (slot_of[i] < 0 || slot_of[i] >= k)
            return set_error(MG_ERR_INVALID_ARG, "%s: pair %d names row %d outside [0,%d)", fn, i, slot_of[i], k);
        if (srcs[i] < 0 || srcs[i] >= g->n)
            return set_error(MG_ERR_INVALID_ARG, "%s: pair %d: source %d outside [0,%d)", fn, i, srcs[i], g->n);
        seen[slot_of[i] >> 5] |= 1u << (slot_of[i] & 31);
    }
    for (int s = 0; s < k; ++s)
        if (!((seen[s >> 5] >> (s & 31)) & 1u)) return set_error(MG_ERR_INVALID_ARG, "%s: row %d has no source", fn, s);
    return MG_OK;
}

// Batched BFS with npairs (row, source) pairs; every row of [0, k) has a source.
static int bfs_batch_run(mg_graph *g, const int32_t *slot_of, const int32_t *srcs, int32_t npairs, int32_t k,
                         int32_t *depth_out, const mg_options &o, mg_stats *stats, const CallTimer &timer) {
    int rc;
    BatchDev *b = nullptr;
    if ((rc = graph_batch_dev(g, false, &b))) return rc;
    auto run = [&]() -> int {
        int rc;
        MG_HIP(hipEventRecord(g->ev0, g->stream));
        if ((rc = bfs_batch_reset(*b, g, g->stream, o.iter_num, batch_alpha(o), batch_beta(o), 0))) return rc;
        if ((rc = bfs_batch_seed_pairs(*b, g, g->stream, slot_of, srcs, npairs))) return rc;
        auto step = [&]() -> int { return bfs_batch_step(*b, g, g->stream); };
        // the step graph reads everything from device memory: one capture per graph
        const int64_t max_steps = std::min<int64_t>((int64_t)o.iter_num, (int64_t)g->n) + 3;
        if ((rc = drive(g, step, max_steps, KIND_BATCH, 0, b->ctl, b->h_ctl))) return rc;
        return finish_batch(g, *b, stats, depth_out, 4 * (size_t)k * g->n, o.out_on_device, timer);
    };
    if ((rc = size_task_queue(g, *b, o))) return rc;
    int regrows = 0;
    while ((rc = run()) == MG_ERR_OVERFLOW && regrows < MAX_REGROWS) {
        ++regrows;
        if ((rc = grow_task_queue(g, *b))) return rc;
        clear_error();
    }
    if (rc == MG_OK && stats) stats->overflow_regrows = regrows;
    return rc;
}

}  // namespace mg

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
    int32_t rows[MG_BATCH_MAX], tsrc[MG_BATCH_MAX];
    for (int s = 0; s < k; ++s) rows[s] = s;
    mg_graph *sg = batch_graph(g, o, srcs, k, tsrc);  // the degree-ordered twin, when the graph keeps one
    if (sg != g && (rc = ensure_workspace(sg))) return rc;
    return bfs_batch_run(sg, rows, sg != g ? tsrc : srcs, k, k, depth_out, o, stats, timer);
}

extern "C" int mg_bfs_batch_sets(mg_graph *g, const int32_t *row_of, const int32_t *srcs, int32_t npairs, int32_t k,
                                 int32_t *depth_out, const mg_options *opt, mg_stats *stats) {
    CallTimer timer;
    mg_options o;
    int rc = begin_search(g, opt, &o);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g->mu);
    if ((rc = check_batch_sets("mg_bfs_batch_sets", g, row_of, srcs, npairs, k))) return rc;
    if (o.directed == 0 && !(g->flags & MG_GRAPH_SYMMETRIC))
        return set_error(MG_ERR_UNSUPPORTED, "mg_bfs_batch_sets: directed=0 needs a symmetric graph");
    std::vector<int32_t> tsrc((size_t)npairs);
    mg_graph *sg = batch_graph(g, o, srcs, npairs, tsrc.data());
    if (sg != g && (rc = ensure_workspace(sg))) return rc;
    return bfs_batch_run(sg, row_of, sg != g ? tsrc.data() : srcs, npairs, k, depth_out, o, stats, timer);
}
