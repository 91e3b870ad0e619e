// Frontier machinery shared by the push-style searches (BFS top-down, SSSP, the
// generic GAS enactor).
//
// One level turns frontier queue q[L%3] into q[(L+1)%3].  Three queues and three
// bitmaps rotate so that, during level L,
//   * bm[L%3]     holds frontier L (read by bottom-up steps),
//   * bm[(L+1)%3] collects frontier L+1 (zero at level start; test-and-set dedupe),
//   * bm[(L+2)%3] still holds frontier L-1 and is cleared (word by word from queue
//     q[(L+2)%3] when that frontier was small, else by a sweep).
// Expansion is tiled: a workgroup takes BLOCK frontier vertices, scans their degrees
// in LDS and walks the flattened edge list edge-parallel; vertices above CHUNK_T
// edges become CHUNK-edge records that the closing kernel spreads over the GPU,
// so a 160K-edge hub costs one level ~160 workgroup-chunks, not one workgroup.
#pragma once

#include "devutil.h"
#include "engine.h"

namespace mg {

struct FrontierArgs {
    Ctl *ctl;
    const int *off[2];   // expansion views: [0] out-edges (CSR), [1] in-edges (CSC) when directed == 0
    const int *adj[2];
    const int *eid[2];   // CSR edge id of each slot of a view (nullptr: the slot is the edge id)
    int nviews;
    uint32_t *bm[3];
    int *q[3];
    ChunkRec *chunks;
    int64_t chunk_cap;
    int n, nwords;
};

__device__ __forceinline__ int ld_dev(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ long long ld_dev(const long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int edge_id(const FrontierArgs &fa, int view, int slot) {
    return fa.eid[view] ? fa.eid[view][slot] : slot;
}

// Degree of v over the expansion views.
__device__ __forceinline__ long long view_degree(const FrontierArgs &fa, int v) {
    long long d = fa.off[0][v + 1] - fa.off[0][v];
    if (fa.nviews > 1) d += fa.off[1][v + 1] - fa.off[1][v];
    return d;
}

// Clears bm[(L+2)%3] (frontier L-1) during level L.
template <int BLOCK>
__device__ void clear_prev(const FrontierArgs &fa, int L, int block, int nblocks) {
    uint32_t *bm = fa.bm[(L + 2) % 3];
    const int *q = fa.q[(L + 2) % 3];
    const int prev = fa.ctl->qsize[(L + 2) % 3];
    const int tid = block * BLOCK + threadIdx.x, nt = nblocks * BLOCK;
    if (prev > fa.nwords / 8) {
        for (int i = tid; i < fa.nwords; i += nt) bm[i] = 0;
    } else {
        for (int i = tid; i < prev; i += nt) bm[q[i] >> 5] = 0;
    }
}

// Marks u for the next frontier (test-and-set on bm[(L+1)%3]) and appends it once.
// Every lane of the wave must call this (wave-aggregated append).
__device__ __forceinline__ bool enqueue_next(const FrontierArgs &fa, int L, int u, bool want) {
    bool first = false;
    if (want) {
        uint32_t *bm = fa.bm[(L + 1) % 3];
        if (!test_bit(bm, u)) first = set_bit_first(bm, u);
    }
    wave_append(first, u, fa.q[(L + 1) % 3], &fa.ctl->qsize[(L + 1) % 3]);
    return first;
}

// Top-down tiles of BLOCK frontier vertices.  P provides
//   int  load(int v)                                   per-vertex payload,
//   void visit(int view, int v, int payload, int slot, int u, bool valid, int L)
// visit is called by every lane of the workgroup in lock-step (valid == false for
// padding lanes) so it may use wave-aggregated appends.  IDENT walks all vertices
// 0..n-1 instead of the queue.
template <int BLOCK, bool CHUNKED, class P, bool IDENT = false>
__device__ void td_tiles(const FrontierArgs &fa, P &pol, int view, int L, int tile0, int tstride) {
    __shared__ int s_beg[BLOCK];
    __shared__ int s_scan[BLOCK];
    __shared__ int s_v[BLOCK];
    __shared__ int s_pay[BLOCK];
    __shared__ int s_wave[BLOCK / WAVE];
    const int *off = fa.off[view], *adj = fa.adj[view];
    const int *qin = fa.q[L % 3];
    const int fin = IDENT ? fa.n : fa.ctl->qsize[L % 3];
    for (int base = tile0 * BLOCK; base < fin; base += tstride * BLOCK) {
        const int i = base + threadIdx.x;
        int v = 0, beg = 0, deg = 0, pay = 0;
        if (i < fin) {
            v = IDENT ? i : qin[i];
            beg = off[v];
            deg = off[v + 1] - beg;
            if (deg > 0) pay = pol.load(v);
        }
        if (CHUNKED && deg > CHUNK_T) {
            const int nch = (deg + CHUNK - 1) / CHUNK;
            const int c0 = atomicAdd(&fa.ctl->nchunks, nch);
            if (c0 + nch <= fa.chunk_cap) {
                for (int c = 0; c < nch; ++c) {
                    ChunkRec r;
                    r.v = v;
                    r.beg = beg + c * CHUNK;
                    r.end = min(beg + (c + 1) * CHUNK, beg + deg);
                    r.pad = view;
                    fa.chunks[c0 + c] = r;
                }
            } else {
                atomicOr(&fa.ctl->changed, 1 << 30);  // capacity is sized so this cannot happen
            }
            deg = 0;
        }
        int total;
        const int ex = block_exclusive_scan<BLOCK>(deg, s_wave, &total);
        s_beg[threadIdx.x] = beg;
        s_scan[threadIdx.x] = ex;
        s_v[threadIdx.x] = v;
        s_pay[threadIdx.x] = pay;
        __syncthreads();
        for (int k0 = 0; k0 < total; k0 += BLOCK) {
            const int k = k0 + threadIdx.x;
            const bool valid = k < total;
            int u = 0, src = 0, p = 0, slot = 0;
            if (valid) {
                int lo = 0, hi = BLOCK - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_scan[mid] <= k) lo = mid;
                    else hi = mid - 1;
                }
                slot = s_beg[lo] + (k - s_scan[lo]);
                u = adj[slot];
                src = s_v[lo];
                p = s_pay[lo];
            }
            pol.visit(view, src, p, slot, u, valid, L);
        }
        __syncthreads();
    }
}

// The edge-chunk records of level L, one workgroup per record at a time.
template <int BLOCK, class P>
__device__ void td_chunks(const FrontierArgs &fa, P &pol, int L) {
    const int nch = (int)min((int64_t)fa.ctl->nchunks, fa.chunk_cap);
    for (int ci = blockIdx.x; ci < nch; ci += gridDim.x) {
        const ChunkRec r = fa.chunks[ci];
        const int *adj = fa.adj[r.pad];
        const int pay = pol.load(r.v);
        for (int e0 = r.beg; e0 < r.end; e0 += BLOCK) {
            const int e = e0 + threadIdx.x;
            const bool valid = e < r.end;
            pol.visit(r.pad, r.v, pay, valid ? e : 0, valid ? adj[e] : 0, valid, L);
        }
    }
}

// Last-workgroup-done: returns true in exactly one thread of the grid, after every
// workgroup's writes are visible.  Must be reached by all threads of every block.
__device__ __forceinline__ bool last_block(Ctl *c) {
    __shared__ int s_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(&c->blocks_done, 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (s_last && threadIdx.x == 0) {
        __threadfence();
        c->blocks_done = 0;
        return true;
    }
    return false;
}

// As last_block, but true in every thread of the last workgroup.
__device__ __forceinline__ bool last_block_all(Ctl *c) {
    __shared__ int s_last_all;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        s_last_all = atomicAdd(&c->blocks_done, 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    const bool last = s_last_all != 0;
    if (last) {
        __threadfence();
        if (threadIdx.x == 0) c->blocks_done = 0;
    }
    return last;
}

// Resets the per-level counters and rotates the queues after level L; returns the
// size of frontier L+1.
__device__ __forceinline__ int rotate_level(Ctl *c, int L) {
    const int nf = ld_dev(&c->qsize[(L + 1) % 3]);
    c->nf_prev = nf;
    c->level = L + 1;
    c->qsize[(L + 2) % 3] = 0;
    c->mf_next = 0;
    c->nchunks = 0;
    c->nhard = 0;
    if (nf == 0 || L + 1 >= c->iter_limit) c->done = 1;
    return nf;
}

// Fills the search-independent fields of the control block.
__device__ __forceinline__ void ctl_reset(Ctl *c, int n, int iter_limit, int frontier0) {
    c->level = 0;
    c->done = iter_limit <= 0 || frontier0 == 0;
    c->dir = DIR_TD;
    c->blocks_done = 0;
    c->qsize[0] = frontier0;
    c->qsize[1] = 0;
    c->qsize[2] = 0;
    c->nchunks = 0;
    c->nhard = 0;
    c->iter_limit = iter_limit;
    c->n = n;
    c->can_bu = 0;
    c->alpha = 14;
    c->beta = 24;
    c->nf_prev = frontier0;
    c->bu_levels = 0;
    c->small_levels = 0;
    c->changed = 0;
    c->mf_next = 0;
    c->m_unexplored = 0;
    c->edges_traversed = 0;
    c->reached = 0;
    c->edges_examined = 0;
    c->dsum = 0;
    c->mf_cur = 0;
}

__global__ void fill_i32(int *p, int v, int64_t n);
__global__ void fill_u32(uint32_t *p, uint32_t v, int64_t n);
__global__ void iota_i32(int *p, int64_t n);

// Sets up the three queues/bitmaps of graph g's workspace and the expansion views.
FrontierArgs make_frontier_args(mg_graph *g, int directed);
int launch_fill(mg_graph *g, int *p, int v, int64_t n);
int launch_fill(mg_graph *g, uint32_t *p, uint32_t v, int64_t n);
int launch_reached_stats(mg_graph *g, const int *vals, int unreached, mg_stats *st);
int finish_search(mg_graph *g, mg_stats *stats, const int *vals, int unreached, void *out, const void *result,
                  size_t bytes, bool out_on_device, const CallTimer &timer);

}  // namespace mg
