// Frontier machinery shared by the push-style searches (BFS top-down, SSSP, CC and
// the generic GAS enactor).
//
// Representation.  Frontier L is always the bitmap bm[L%3] (n bits: 512 KB at 4M
// vertices, L2-resident).  Three bitmaps rotate: during level L, bm[L%3] is read,
// bm[(L+1)%3] collects frontier L+1 (zero at level start; test-and-set dedupe) and
// bm[(L+2)%3] (frontier L-1) is cleared.  The single-workgroup small-frontier kernel
// additionally keeps frontier queues (q[L%3]) so a level costs O(frontier), not
// O(n); multi-workgroup levels never append to a global queue.
//
// Contention.  A same-address atomic costs ~10 ns on the device's memory side and
// they serialise, so nothing here issues one per wave: counters are per-workgroup
// partial slots summed by one thread, queue/chunk reservations take one atomic per
// workgroup round, and the level is closed by the next step's single-workgroup
// kernel instead of a grid-wide last-block counter.
//
// Expansion.  A workgroup compacts the set bits of a slice of the frontier bitmap
// into an LDS vertex list, then walks it in tiles of BLOCK vertices: degrees are
// scanned in LDS and the flattened edge list is processed edge-parallel; vertices
// above CHUNK_T edges become CHUNK-edge records that the second level kernel
// spreads over the whole GPU (a 160K-edge hub becomes ~160 workgroup-chunks).
#pragma once

#include <type_traits>
#include <utility>

#include "devutil.h"
#include "engine.h"

namespace mg {

struct FrontierArgs {
    Ctl *ctl;
    // Views: [0] out-edges (CSR); [1] in-edges (CSC) expanded as well when
    // directed == 0; [2] the in-edges of the traversed graph for pull steps (the CSR
    // itself when the graph is symmetric, else the CSC; off[2] == nullptr when none).
    const int *off[3];
    const int *adj[3];
    const int *eid[3];   // CSR edge id of each slot of a view (nullptr: the slot is the edge id)
    int nviews;          // views expanded by push steps (1 or 2)
    uint32_t *bm[3];
    int *q[3];
    ChunkRec *chunks;
    int64_t chunk_cap;
    LevelPart *part;     // per-workgroup partial counters, 3 * BIG_GRID slots
    int n, nwords;
    long long *trace;    // per-level trace (TRACE_CAP levels), written by the level close
};

// Records level L's outcome: the next frontier's size and out-edges and the
// direction the level ran in (one thread).
__device__ __forceinline__ void trace_level(const FrontierArgs &fa, int L, int dir, long long nf, long long mf) {
    if (fa.trace && L < TRACE_CAP) {
        fa.trace[2 * L] = nf | ((long long)dir << 40);
        fa.trace[2 * L + 1] = mf;
    }
}

// Workgroup-scope access for single-workgroup loops: every reader and writer of the
// data during the launch is in the one workgroup (one CU), so the CU's caches are
// coherent enough and nothing needs the device-wide path (other kernels see the
// results after the launch ends).
__device__ __forceinline__ int ld_wg(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void min_wg(int *p, int v) {
    __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ bool set_bit_first_wg(uint32_t *bm, int v) {
    const uint32_t bit = 1u << (v & 31);
    return !(__hip_atomic_fetch_or(&bm[v >> 5], bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & bit);
}
__device__ __forceinline__ void clear_bit_wg(uint32_t *bm, int v) {
    __hip_atomic_fetch_and(&bm[v >> 5], ~(1u << (v & 31)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

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

// Deterministic workgroup sum (wave butterfly, then waves in order); every thread
// gets the result.
template <int BLOCK, class T>
__device__ T block_sum(T x) {
    __shared__ T s_w[BLOCK / WAVE];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    __syncthreads();
    if (lane_id() == 0) s_w[threadIdx.x / WAVE] = x;
    __syncthreads();
    T t = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / WAVE; ++w) t += s_w[w];
    return t;
}

// Workgroup minimum; every thread gets the result.
template <int BLOCK, class T>
__device__ T block_min(T x) {
    __shared__ T s_m[BLOCK / WAVE];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const T y = __shfl_xor(x, o);
        x = y < x ? y : x;
    }
    __syncthreads();
    if (lane_id() == 0) s_m[threadIdx.x / WAVE] = x;
    __syncthreads();
    T t = s_m[0];
#pragma unroll
    for (int w = 1; w < BLOCK / WAVE; ++w) t = s_m[w] < t ? s_m[w] : t;
    return t;
}

// Level counters: every workgroup of a level's kernels adds its sums into slot
// blockIdx % PART_GROUPS (skipped when all are zero); the next close reads the slots
// and clears them, so one workgroup reads PART_GROUPS slots rather than one per
// workgroup of the level (a few thousand, written on every XCD).  The kernel that
// starts a search clears them.
constexpr int PART_GROUPS = 256;
// The slot's fourth field carries a minimum (the smallest pending distance of a
// bucketed SSSP level) as the key PART_MIN_BASE - value under atomicMax, so a cleared
// slot (0) means "none" for every user.
constexpr long long PART_MIN_BASE = 1LL << 62;
template <int BLOCK>
__device__ __forceinline__ void write_part(const FrontierArgs &fa, long long nf, long long mf, long long ex,
                                           long long mn = LLONG_MAX) {
    nf = block_sum<BLOCK>(nf);
    mf = block_sum<BLOCK>(mf);
    ex = block_sum<BLOCK>(ex);
    mn = block_min<BLOCK>(mn);
    if (threadIdx.x == 0) {
        LevelPart *p = &fa.part[blockIdx.x % PART_GROUPS];
        if (nf) atomicAdd((unsigned long long *)&p->nf, (unsigned long long)nf);
        if (mf) atomicAdd((unsigned long long *)&p->mf, (unsigned long long)mf);
        if (ex) atomicAdd((unsigned long long *)&p->ex, (unsigned long long)ex);
        if (mn < PART_MIN_BASE)
            atomicMax((unsigned long long *)&p->pad, (unsigned long long)(PART_MIN_BASE - mn));
    }
}

// Thread t < PART_GROUPS of workgroup 0 of a search's first kernel clears slot t.
__device__ __forceinline__ void clear_parts(const FrontierArgs &fa) {
    if (threadIdx.x < PART_GROUPS) fa.part[threadIdx.x] = LevelPart{0, 0, 0, 0};
}

// Sums the level counters (whole workgroup, BLOCK >= PART_GROUPS) and clears them.
// The minimum written with write_part comes back in *mn (LLONG_MAX when none).
template <int BLOCK>
__device__ __forceinline__ void take_level_parts(const FrontierArgs &fa, long long &nf, long long &mf,
                                                 long long &ex, long long *mn = nullptr) {
    static_assert(BLOCK >= PART_GROUPS, "one slot per thread");
    nf = mf = ex = 0;
    long long m = LLONG_MAX;
    if (threadIdx.x < PART_GROUPS) {
        const LevelPart p = fa.part[threadIdx.x];
        nf = p.nf;
        mf = p.mf;
        ex = p.ex;
        if (p.pad) m = PART_MIN_BASE - p.pad;
        fa.part[threadIdx.x] = LevelPart{0, 0, 0, 0};
    }
    nf = block_sum<BLOCK>(nf);
    mf = block_sum<BLOCK>(mf);
    ex = block_sum<BLOCK>(ex);
    if (mn) *mn = block_min<BLOCK>(m);
}

// Clears bm[(L+2)%3] (frontier L-1) during level L by a sweep of 16-byte stores
// (bitmaps are 256-byte aligned; nwords is even).
__device__ __forceinline__ void clear_prev_sweep(const FrontierArgs &fa, int L, int tid, int nt) {
    uint32_t *bm = fa.bm[(L + 2) % 3];
    const int n4 = fa.nwords / 4;
    uint4 *p = reinterpret_cast<uint4 *>(bm);
    for (int i = tid; i < n4; i += nt) p[i] = make_uint4(0, 0, 0, 0);
    for (int i = 4 * n4 + tid; i < fa.nwords; i += nt) bm[i] = 0;
}

// Marks u in the next-frontier bitmap; true for the one call that set the bit.
__device__ __forceinline__ bool mark_next(const FrontierArgs &fa, int L, int u) {
    uint32_t *bm = fa.bm[(L + 1) % 3];
    if (test_bit(bm, u)) return false;
    return set_bit_first(bm, u);
}

// Sets the bits of bitmap bm for the targets u[k] flagged in `need` (bit k) and,
// for each bit this lane set first, calls on_first(u) and counts it in nf with its
// out-degree in mf; mirror (if not null) gets the same newly set bits.  u[k] must
// be a vertex for every k.  The lane's test-and-sets go out together, then the
// degree loads (clamped to vertex 0's shared line where not needed): old[k] is left
// unassigned where no atomic is issued (it is read only under the same flag), so no
// merge of the atomic's result waits for it inside the branch.
// RUNS (for dense claims such as a BFS level from a hub): with the whole wave active,
// lanes whose targets share a bitmap word in a run of consecutive lanes (sorted lists)
// OR their bits together and the run's last lane issues one atomic for the run; each
// lane then reads its bit in the returned word (and in the bits of the lanes before
// it, for repeated targets).  Where claims are sparse (fewer than RUNS_MIN lanes with
// one) the per-lane form is cheaper.
constexpr int RUNS_MIN = 16;  // lanes with claims below which a wave takes the per-lane form
template <int K, bool RUNS, class F>
__device__ __forceinline__ void claim_batch(const FrontierArgs &fa, uint32_t *bm, const int (&u)[K], uint32_t need,
                                            long long &nf, long long &mf, F on_first, uint32_t *mirror = nullptr) {
    if (!wave_ballot(need != 0)) return;
    bool first[K];
    if (RUNS && wave_ballot(true) == ~0ull && __popcll(wave_ballot(need != 0)) >= RUNS_MIN) {
        const int lane = lane_id();
        const uint64_t upto = lane == WAVE - 1 ? ~0ull : ((2ull << lane) - 1);  // lanes 0..lane
        uint32_t mine[K], incl[K], prev[K], old[K];
        int tail[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            mine[k] = ((need >> k) & 1u) << (u[k] & 31);
            const int w = u[k] >> 5;
            const int wp = __shfl_up(w, 1);
            const uint64_t heads = wave_ballot(lane == 0 || wp != w);
            const int start = WAVE - 1 - __builtin_clzll(heads & upto);
            const uint64_t later = heads & ~upto;
            tail[k] = later ? __builtin_ctzll(later) - 1 : WAVE - 1;
            incl[k] = seg_or_scan(mine[k], lane, start);
            const uint32_t pv = __shfl_up(incl[k], 1);
            prev[k] = lane > start ? pv : 0u;
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            old[k] = 0;  // read by __shfl below in every lane: defined everywhere
            if (lane == tail[k] && incl[k]) old[k] = atomicOr(&bm[u[k] >> 5], incl[k]);
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t o = __shfl(old[k], tail[k]);
            first[k] = mine[k] && !(o & mine[k]) && !(prev[k] & mine[k]);
            if (mirror && lane == tail[k] && incl[k]) {
                const uint32_t fresh = incl[k] & ~o;  // bits this run set first
                if (fresh) atomicOr(&mirror[u[k] >> 5], fresh);
            }
        }
    } else {
        uint32_t old[K];
#pragma unroll
        for (int k = 0; k < K; ++k)
            if ((need >> k) & 1u) old[k] = atomicOr(&bm[u[k] >> 5], 1u << (u[k] & 31));
#pragma unroll
        for (int k = 0; k < K; ++k) {
            first[k] = ((need >> k) & 1u) ? !((old[k] >> (u[k] & 31)) & 1u) : false;
            if (mirror && first[k]) atomicOr(&mirror[u[k] >> 5], 1u << (u[k] & 31));
        }
    }
    int o0[K], o1[K], p0[K], p1[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int d = first[k] ? u[k] : 0;
        o0[k] = fa.off[0][d];
        o1[k] = fa.off[0][d + 1];
        if (fa.nviews > 1) {
            p0[k] = fa.off[1][d];
            p1[k] = fa.off[1][d + 1];
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (first[k]) {
            on_first(u[k]);
            ++nf;
            mf += o1[k] - o0[k];
            if (fa.nviews > 1) mf += p1[k] - p0[k];
        }
}

// One tile of up to BLOCK frontier vertices (thread i holds vertex v_i, or -1).
// P provides
//   int  load(int v)                                  per-vertex payload,
//   bool visit(int view, int v, int payload, int slot, int u, bool valid, int L)
// visit is called by every thread of the workgroup in lock-step (valid == false for
// padding) and returns true when u newly entered the next frontier.  With CHUNKED,
// vertices above CHUNK_T edges become chunk records; otherwise they are expanded in
// the tile.  `emit(u, newly)` is called, also in lock-step, after each visit.
// As expand_tile, over the explicit slot range [beg, beg + deg) of vertex v.
template <int BLOCK, bool CHUNKED, class P, class Emit>
__device__ void expand_tile_range(const FrontierArgs &fa, P &pol, int view, int L, int v, int beg, int deg,
                                  Emit emit);

template <int BLOCK, bool CHUNKED, class P, class Emit>
__device__ void expand_tile(const FrontierArgs &fa, P &pol, int view, int L, int v, Emit emit) {
    const int *off = fa.off[view];
    int beg = 0, deg = 0;
    if (v >= 0) {
        beg = off[v];
        deg = off[v + 1] - beg;
    }
    expand_tile_range<BLOCK, CHUNKED>(fa, pol, view, L, v, beg, deg, emit);
}

template <int BLOCK, bool CHUNKED, class P, class Emit>
__device__ void expand_tile_range(const FrontierArgs &fa, P &pol, int view, int L, int v, int beg, int deg,
                                  Emit emit) {
    __shared__ int s_beg[BLOCK];
    __shared__ int s_scan[BLOCK];
    __shared__ int s_v[BLOCK];
    __shared__ int s_pay[BLOCK];
    __shared__ int s_wave[BLOCK / WAVE];
    __shared__ int s_cbase;
    const int *adj = fa.adj[view];
    int pay = 0;
    if (v < 0) deg = 0;
    if (deg > 0) pay = pol.load(v);
    if (CHUNKED) {
        const int nch = deg > CHUNK_T ? (deg + CHUNK - 1) / CHUNK : 0;
        int ctot;
        const int cex = block_exclusive_scan<BLOCK>(nch, s_wave, &ctot);
        if (ctot > 0) {
            if (threadIdx.x == 0) s_cbase = atomicAdd(&fa.ctl->nchunks, ctot);
            __syncthreads();
            const int c0 = s_cbase + cex;
            if (nch && c0 + nch <= fa.chunk_cap) {
                for (int c = 0; c < nch; ++c) {
                    ChunkRec r;
                    r.v = v;
                    r.beg = beg + c * CHUNK;
                    r.end = min(beg + (c + 1) * CHUNK, beg + deg);
                    r.pad = view;
                    fa.chunks[c0 + c] = r;
                }
            } else if (nch) {
                atomicOr(&fa.ctl->changed, 1 << 30);  // capacity is sized so this cannot happen
            }
        }
        if (nch) deg = 0;
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
        const bool newly = pol.visit(view, src, p, slot, u, valid, L);
        emit(u, newly);
    }
    __syncthreads();
}

// Multi-workgroup top-down over the frontier bitmap `fr`: each round a workgroup
// compacts BLOCK/2 words (BLOCK*16 vertices) into LDS and expands them in tiles.
template <int BLOCK, class P>
__device__ void expand_bitmap(const FrontierArgs &fa, P &pol, int view, int L, const uint32_t *fr) {
    constexpr int WPR = BLOCK / 2;
    __shared__ int s_q[WPR * 32];
    __shared__ int s_wave2[BLOCK / WAVE];
    auto no_emit = [](int, bool) {};
    for (int w0 = blockIdx.x * WPR; w0 < fa.nwords; w0 += gridDim.x * WPR) {
        const int w = w0 + (int)threadIdx.x;
        uint32_t word = (threadIdx.x < WPR && w < fa.nwords) ? fr[w] : 0u;
        int total;
        int k = block_exclusive_scan<BLOCK>(__popc(word), s_wave2, &total);
        while (word) {
            const int b = __ffs(word) - 1;
            s_q[k++] = w * 32 + b;
            word &= word - 1;
        }
        __syncthreads();
        for (int t0 = 0; t0 < total; t0 += BLOCK) {
            const int i = t0 + (int)threadIdx.x;
            const int v = i < total ? s_q[i] : -1;
            expand_tile<BLOCK, true>(fa, pol, view, L, v, no_emit);
        }
        __syncthreads();
    }
}

// Policies that postpone frontier vertices define bool defer(int v, int L): called
// for every frontier vertex with edges in the view (once per expanded view) before
// they are expanded; true skips the vertex (the policy carries it to the next
// frontier itself).
template <class P, class = void>
struct policy_defers : std::false_type {};
template <class P>
struct policy_defers<P, std::void_t<decltype(std::declval<P &>().defer(0, 0))>> : std::true_type {};

// Multi-workgroup push over the frontier bitmap `fr` without workgroup barriers in
// the edge loop: each round a workgroup compacts BLOCK/2 bitmap words into LDS; a
// lane walks the edges of each listed vertex of degree <= LANE_T itself, longer
// lists become TASK-edge task records (one reservation atomic per workgroup round)
// for expand_tasks.  The policy's visit_range(view, v, payload, beg, end, stride, L)
// handles slots beg, beg+stride, ... < end of one lane; it must not use wave-wide
// operations (lanes diverge) and should batch its loads (see load4).
template <int BLOCK, class P>
__device__ void expand_lanes(const FrontierArgs &fa, P &pol, int view, int L, const uint32_t *fr) {
    constexpr int WPR = BLOCK / 2;
    __shared__ int s_q[WPR * 32];
    __shared__ int s_wave2[BLOCK / WAVE];
    __shared__ int s_tbase;
    __shared__ int s_tex[BLOCK], s_tv[BLOCK], s_tbeg[BLOCK], s_tend[BLOCK];
    const int *off = fa.off[view];
    for (int w0 = blockIdx.x * WPR; w0 < fa.nwords; w0 += gridDim.x * WPR) {
        const int w = w0 + (int)threadIdx.x;
        uint32_t word = (threadIdx.x < WPR && w < fa.nwords) ? fr[w] : 0u;
        int total;
        int k = block_exclusive_scan<BLOCK>(__popc(word), s_wave2, &total);
        while (word) {
            const int b = __ffs(word) - 1;
            s_q[k++] = w * 32 + b;
            word &= word - 1;
        }
        __syncthreads();
        for (int t0 = 0; t0 < total; t0 += BLOCK) {
            const int i = t0 + (int)threadIdx.x;
            int v = 0, beg = 0, deg = 0;
            if (i < total) {
                v = s_q[i];
                beg = off[v];
                deg = off[v + 1] - beg;
                if constexpr (policy_defers<P>::value)
                    if (deg > 0 && pol.defer(v, L)) deg = 0;
            }
            const int nt = deg > LANE_T ? (deg + TASK - 1) / TASK : 0;
            int ttot;
            const int tex = block_exclusive_scan<BLOCK>(nt, s_wave2, &ttot);
            if (ttot > 0) {
                // the workgroup writes the records together (a hub has hundreds): record j
                // belongs to the last thread whose exclusive offset is <= j
                if (threadIdx.x == 0) s_tbase = atomicAdd(&fa.ctl->nchunks, ttot);
                s_tex[threadIdx.x] = tex;
                s_tv[threadIdx.x] = v;
                s_tbeg[threadIdx.x] = beg;
                s_tend[threadIdx.x] = beg + deg;
                __syncthreads();
                for (int j = threadIdx.x; j < ttot; j += BLOCK) {
                    int o = 0;
#pragma unroll
                    for (int st = BLOCK / 2; st > 0; st >>= 1)
                        if (s_tex[o + st] <= j) o += st;
                    const int idx = s_tbase + j;
                    if (idx < fa.chunk_cap) {
                        ChunkRec r;
                        r.v = s_tv[o];
                        r.beg = s_tbeg[o] + (j - s_tex[o]) * TASK;
                        r.end = min(r.beg + TASK, s_tend[o]);
                        r.pad = view;
                        fa.chunks[idx] = r;
                    } else {
                        atomicOr(&fa.ctl->changed, 1 << 30);  // capacity is sized so this cannot happen
                    }
                }
                __syncthreads();
            }
            if (nt == 0 && deg > 0) pol.visit_range(view, v, pol.load(v), beg, beg + deg, 1, L);
        }
        __syncthreads();
    }
}

// The task records of the level, one wave per record, lanes striding its edges.
template <int BLOCK, class P>
__device__ void expand_tasks(const FrontierArgs &fa, P &pol, int L) {
    const int nch = (int)min((int64_t)fa.ctl->nchunks, fa.chunk_cap);
    const int lane = lane_id();
    const int wave = (blockIdx.x * BLOCK + (int)threadIdx.x) / WAVE, nwaves = gridDim.x * BLOCK / WAVE;
    for (int t = wave; t < nch; t += nwaves) {
        const ChunkRec r = fa.chunks[t];
        pol.visit_range(r.pad, r.v, pol.load(r.v), r.beg + lane, r.end, WAVE, L);
    }
}

// Edge slots beg, beg+stride, ... < end of `view` in groups of 4 with the neighbour
// ids loaded together: out[k] = adj[slot_k] (or -1 past the end).
__device__ __forceinline__ void load4(const int *adj, int e, int end, int stride, int (&u)[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) u[k] = load_or(adj, e + k * stride, e + k * stride < end, e, -1);  // e < end
}

// The edge-chunk records of the level, one workgroup per record at a time.
template <int BLOCK, class P>
__device__ void expand_chunks(const FrontierArgs &fa, P &pol, int L) {
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

// Builds queue q[L%3] of frontier L from its bitmap inside one workgroup (used when
// the small-frontier kernel takes over from multi-workgroup levels).  Each thread
// takes 16 consecutive words (four 16-byte loads in flight) per round, so a 4M-vertex
// bitmap is 8 rounds of one workgroup rather than 128 load-then-barrier rounds.
template <int BLOCK>
__device__ void queue_from_bitmap_to(const FrontierArgs &fa, int L, int *q, int *s_count, int cap = INT_MAX) {
    constexpr int WPT = 16;
    __shared__ int s_wave3[BLOCK / WAVE];
    const uint32_t *bm = fa.bm[L % 3];
    for (int w0 = 0; w0 < fa.nwords; w0 += BLOCK * WPT) {
        const int wb = w0 + (int)threadIdx.x * WPT;
        uint32_t words[WPT];
        if (wb + WPT <= fa.nwords) {
            const uint4 *p = reinterpret_cast<const uint4 *>(bm + wb);
#pragma unroll
            for (int j = 0; j < WPT / 4; ++j) {
                const uint4 x = p[j];
                words[4 * j] = x.x;
                words[4 * j + 1] = x.y;
                words[4 * j + 2] = x.z;
                words[4 * j + 3] = x.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < WPT; ++j) words[j] = wb + j < fa.nwords ? bm[wb + j] : 0u;
        }
        int c = 0;
#pragma unroll
        for (int j = 0; j < WPT; ++j) c += __popc(words[j]);
        int total;
        int k = block_exclusive_scan<BLOCK>(c, s_wave3, &total) + *s_count;
#pragma unroll
        for (int j = 0; j < WPT; ++j) {
            uint32_t word = words[j];
            while (word) {
                const int b = __ffs(word) - 1;
                if (k < cap) q[k] = (wb + j) * 32 + b;
                ++k;
                word &= word - 1;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) *s_count += total;
        __syncthreads();
    }
}

// As queue_from_bitmap_to, into the global queue q[L % 3].
template <int BLOCK>
__device__ void queue_from_bitmap(const FrontierArgs &fa, int L, int *s_count) {
    queue_from_bitmap_to<BLOCK>(fa, L, fa.q[L % 3], s_count);
}

// Fills the search-independent fields of the control block.
__device__ __forceinline__ void ctl_reset(Ctl *c, int n, int iter_limit, int frontier0) {
    c->level = 0;
    c->done = iter_limit <= 0 || frontier0 == 0;
    c->dir = DIR_TD;
    c->steps = 0;
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
    c->pending = 0;
    c->qvalid = 1;
    c->mf_next = 0;
    c->m_unexplored = 0;
    c->edges_traversed = 0;
    c->reached = 0;
    c->edges_examined = 0;
    c->dsum = 0;
    c->mf_cur = 0;
}

// Advances the control block past level L given the next frontier's size nf and
// out-edge count mf: rotation, per-level counters, termination.
__device__ __forceinline__ void advance_level(Ctl *c, int L, long long nf, long long mf) {
    c->nf_prev = (int)nf;
    c->mf_cur = mf;
    c->level = L + 1;
    c->qsize[(L + 1) % 3] = (int)nf;
    c->qsize[(L + 2) % 3] = 0;
    c->mf_next = 0;
    c->nchunks = 0;
    c->nhard = 0;
    if (nf == 0 || L + 1 >= c->iter_limit) c->done = 1;
}

__global__ void fill_i32(int *p, int v, int64_t n);
__global__ void fill_u32(uint32_t *p, uint32_t v, int64_t n);
__global__ void iota_i32(int *p, int64_t n);

// Sets up the bitmaps/queues of graph g's workspace and the expansion views.
FrontierArgs make_frontier_args(mg_graph *g, int directed);
// out[v] = vals[perm[v]] for v < n on st: a twin's row in the caller's numbering (perm
// maps caller ids to twin ids); out may be device memory or pinned host memory.
int launch_unpermute(const int *perm, const void *vals, int n, void *out, hipStream_t st);
int launch_fill(mg_graph *g, int *p, int v, int64_t n);
int launch_fill(mg_graph *g, uint32_t *p, uint32_t v, int64_t n);
int ensure_pull_tasks(mg_graph *g, const int *d_in_off);
// Tasks of the dense SSSP pull levels: {v, beg, end} = slots [beg, end) of the
// in-list of a vertex v with more than DENSE_BIG in-edges (DENSE_T per piece), or
// {-1, va, vb} = the vertices [va, vb) (va a multiple of 32, vb too or n, at most
// DENSE_V of them) whose in-lists are at most DENSE_BIG long, about DENSE_T edges in all.
int ensure_dense_tasks(mg_graph *g, const int *d_in_off);
// Prints the per-level trace of the search that just finished to stderr (the
// reference's verbose per-iteration frontier sizes).
int print_trace(mg_graph *g, const char *what, const char *dir0, const char *dir1);
// Queues what follows a finished search: the result back to the caller's ids (a twin's
// search), the closing event and, with_stats, the TEPS count over vals.
int queue_finish(mg_graph *g, bool with_stats, const int *vals, int unreached, const void *result, size_t bytes);
int finish_search(mg_graph *g, mg_stats *stats, const int *vals, int unreached, void *out, const void *result,
                  size_t bytes, bool out_on_device, const CallTimer &timer, bool tail_queued = false);

}  // namespace mg
