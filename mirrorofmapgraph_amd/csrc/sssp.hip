// Single-source shortest paths: the engine's specialised path for the sssp program
// (Algorithms/SSSP/sssp.h).
//
// The reference program is push-style frontier Bellman-Ford: expand emits
// dist[v] + w for out-edges that improve on dist[nbr] (sssp.h:342-357), contract
// atomicMin's it into gather_tmp (sssp.h:384-401), apply/post_apply commit the
// minimum and mark changed vertices (sssp.h:242-280), and changed vertices form the
// next frontier.  Here each level is one of two fused steps, chosen on the device:
//   * push  (sparse frontiers) relaxes the frontier's out-edges straight into dist
//           with atomicMin and marks improved targets in the next-frontier bitmap;
//   * pull  (dense frontiers, when an in-edge view exists) lets every vertex take
//           the minimum of dist[u] + w over all its in-edges (see dense_pull):
//           edges streamed once, the minimum reduced in registers and LDS.
// Both are Bellman-Ford relaxations; any vertex whose distance drops during a level
// is in the next frontier, so the fixed point is the same.  Integer shortest
// distances are unique, so the result equals the reference's: dist[src] = 0,
// unreached = INT_MAX (sssp.h:42-43, 95-100).  The step structure is BFS's: one
// single-workgroup kernel for small frontiers and the level close, two whole-GPU
// kernels per large level.
#include <climits>

#include "frontier.h"

namespace mg {

enum { DIR_PUSH = DIR_TD, DIR_PULL = DIR_BU };

struct SsspArgs {
    FrontierArgs fa;
    int *dist;
    const int *w;
    int G;                 // workgroups of the whole-GPU level kernels
    // pull levels: the dense tasks, all in one launch
    const ChunkRec *dpull;
    int ndpull;
    int push_range;  // symmetric graph with dense tasks: wide push levels use range_push
    // buckets: width, far list of the small-frontier loop (n entries) and its bitmap
    long long delta;
    int *far;
    uint32_t *farbm;
    // bucketed whole-GPU levels: a frontier vertex at or above ctl->thr is carried to
    // the next frontier unrelaxed; once the pending vertices' out-edges fall below
    // bucket_edges the bound is lifted (UNBUCKETED) and the tail runs as Bellman-Ford
    int bucketed;
    long long bucket_edges;
    int wmin;  // smallest edge weight (>= 0)
};

constexpr long long UNBUCKETED = LLONG_MAX / 4;  // a bucket bound no distance reaches

__device__ __forceinline__ long long edge_w(const SsspArgs &a, int view, int slot) {
    return a.w[edge_id(a.fa, view, slot)];
}

// Push visit: relax (v, u); mark u when it improved.
// Lowers dist[u[k]] to nd[k] where that improves it and marks the improved targets
// in the next frontier nx (claim_batch: the lane's atomics go out together).  u[k]
// must be a vertex even where ok[k] is false (callers load from clamped slots).  A
// stale (larger) du[k] only causes a redundant relaxation: the target's distance
// dropped this level either way, so it belongs to the next frontier.
// mn takes the smallest improved distance (the next frontier's pending minimum, for
// the bucket bound).
template <int K>
__device__ __forceinline__ void relax_k(const SsspArgs &a, uint32_t *nx, const int (&u)[K],
                                        const long long (&nd)[K], const int (&du)[K], const bool (&ok)[K],
                                        long long &nf, long long &mf, long long &mn) {
    uint32_t need = 0;
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (ok[k] && nd[k] < (long long)du[k]) {
            need |= 1u << k;
            mn = nd[k] < mn ? nd[k] : mn;
        }
#pragma unroll
    for (int k = 0; k < K; ++k)
        if ((need >> k) & 1u) atomicMin(&a.dist[u[k]], (int)nd[k]);
    claim_batch<K, true>(a.fa, nx, u, need, nf, mf, [](int) {});
}

struct SsspPush {
    const SsspArgs &a;
    long long nf = 0, mf = 0, ex = 0;
    long long mn = LLONG_MAX;  // smallest distance entering the next frontier
    long long thr;             // bucket bound of this level (UNBUCKETED: relax every frontier vertex)
    __device__ explicit SsspPush(const SsspArgs &args)
        : a(args), thr(args.bucketed ? args.fa.ctl->thr : UNBUCKETED) {}
    // device-coherent: the small-frontier loop reads distances lowered earlier in the launch
    __device__ int load(int v) const { return ld_dev(&a.dist[v]); }
    // Bucketed levels: a frontier vertex at or above the bound waits in the next
    // frontier, unrelaxed.  A stale (larger) distance read here only postpones a vertex
    // that stays pending; the one that is relaxed reads its distance again (load).
    __device__ bool defer(int v, int L) {
        if (thr >= UNBUCKETED) return false;
        const long long dv = a.dist[v];
        if (dv < thr) return false;
        mn = dv < mn ? dv : mn;
        if (set_bit_first(a.fa.bm[(L + 1) % 3], v)) {
            ++nf;
            mf += view_degree(a.fa, v);
        }
        return true;
    }
    // u's distance dropped (to at most nd) this level: mark it once.  A stale
    // (larger) dist[u] read only causes a redundant relaxation: dist[u] dropped
    // this level either way, so u belongs to the next frontier.
    __device__ bool relax(long long nd, int du, int u, int L) {
        if (nd >= (long long)du) return false;
        mn = nd < mn ? nd : mn;
        atomicMin(&a.dist[u], (int)nd);
        if (!mark_next(a.fa, L, u)) return false;
        ++nf;
        mf += view_degree(a.fa, u);
        return true;
    }
    __device__ bool visit(int view, int, int d, int slot, int u, bool valid, int L) {
        if (!valid) return false;
        ++ex;
        return relax((long long)d + edge_w(a, view, slot), ld_dev(&a.dist[u]), u, L);
    }
    // four targets, weights and target distances in flight together, then relax_k
    __device__ void visit_range(int view, int, int d, int beg, int end, int stride, int L) {
        const int *adj = a.fa.adj[view];
        const int *eid = a.fa.eid[view];
        uint32_t *nx = a.fa.bm[(L + 1) % 3];
        for (int e = beg; e < end; e += 4 * stride) {
            int sl[4], u[4], wt[4], du[4];
            long long nd[4];
            bool ok[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                ok[k] = e + k * stride < end;
                sl[k] = ok[k] ? e + k * stride : e;  // e < end: in bounds
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) u[k] = adj[sl[k]];
            if (eid) {
                int ei[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) ei[k] = eid[sl[k]];
#pragma unroll
                for (int k = 0; k < 4; ++k) wt[k] = a.w[ei[k]];
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) wt[k] = a.w[sl[k]];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) du[k] = a.dist[u[k]];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                nd[k] = (long long)d + wt[k];
                ex += ok[k];
            }
            relax_k<4>(a, nx, u, nd, du, ok, nf, mf, mn);
        }
    }
};

// Pull: the best dist[u] + w over the in-edge slots [beg, end) of vertex v, PULL_U
// loads in flight (neighbour ids and weights, then the distances they need).  No
// frontier test: a vertex u outside the frontier has relaxed its out-edges with its
// present distance already (its last change put it in a frontier), so its candidate
// dist[u] + w cannot beat dist[v]; unreached u give INT_MAX + w.
//
// Bound: every frontier vertex has dist >= fmin (the level's lower bound), so an edge
// of weight w can only improve v when fmin + w < dist[v]; the distance gather of any
// other edge is skipped (its lane reads dist[hot] instead, one cached line shared by
// the wave, and contributes nothing).  On R-MAT from a hub most distances are small
// after the first wide level, so the dense levels gather for a few percent of their
// edges: the pull streams neighbour ids and weights and gathers little.
constexpr int PULL_U = 4;
template <bool EID>
__device__ __forceinline__ long long pull_best(const SsspArgs &a, int beg, int end, int step, long long fmin,
                                               long long dv, int hot) {
    long long best = dv;  // candidates at or above the current distance change nothing
    const int *adj = a.fa.adj[2];
    const int *eid = a.fa.eid[2];
    for (int e = beg; e < end; e += PULL_U * step) {
        int sl[PULL_U], u[PULL_U], wt[PULL_U];
        bool ok[PULL_U];
#pragma unroll
        for (int k = 0; k < PULL_U; ++k) {
            ok[k] = e + k * step < end;
            sl[k] = ok[k] ? e + k * step : e;  // e < end: in bounds
        }
#pragma unroll
        for (int k = 0; k < PULL_U; ++k) u[k] = adj[sl[k]];
        if (EID) {
            int ei[PULL_U];
#pragma unroll
            for (int k = 0; k < PULL_U; ++k) ei[k] = eid[sl[k]];
#pragma unroll
            for (int k = 0; k < PULL_U; ++k) wt[k] = a.w[ei[k]];
        } else {
#pragma unroll
            for (int k = 0; k < PULL_U; ++k) wt[k] = a.w[sl[k]];
        }
        bool need[PULL_U];
        int du[PULL_U];
#pragma unroll
        for (int k = 0; k < PULL_U; ++k) need[k] = ok[k] && fmin + wt[k] < best;
#pragma unroll
        for (int k = 0; k < PULL_U; ++k) du[k] = a.dist[need[k] ? u[k] : hot];
#pragma unroll
        for (int k = 0; k < PULL_U; ++k) {
            const long long cand = need[k] ? (long long)du[k] + wt[k] : LLONG_MAX;
            best = cand < best ? cand : best;
        }
    }
    return best;
}

// Dense pull level: every vertex takes min(dist[u] + w) over all its in-edges (no
// frontier test, see pull_best; the same fmin bound skips the distance gathers of
// edges that cannot improve), in one launch over the static dense tasks:
//   * a piece of a long list: a wave minimum, one atomicMin into dist[v];
//   * a range of up to DENSE_V vertices of whole bitmap words with short lists: the
//     wave walks the concatenation of their in-lists in windows of 64 slots, one per
//     lane, with the range index below (so neighbour ids and weights are coalesced
//     loads and the neighbour distances the only gathers), DENSE_WIN windows in flight
//     together.  Owners are non-decreasing along a window, so a segmented minimum scan
//     across the wave reduces each owner's run and its last lane folds it into the
//     owner's LDS minimum (which starts at the owner's distance and bounds its
//     gathers); the range's own vertices are then updated with plain stores (no other
//     task writes them) and their frontier words set with one atomicOr each
//     (long-list vertices of the same words are marked by their pieces).
// mn takes the smallest improved distance (the next level's fmin).
constexpr int DENSE_WIN = 2;  // windows in flight (4 and 8: more registers, fewer waves, slower)
constexpr int DENSE_PER = DENSE_V / WAVE;  // vertices per lane in a range task
constexpr int DENSE_DIRECT = 16;           // candidates a window folds in with LDS atomics

// Range index: a range task's lists concatenated, for the wave that walks them.  Bit p
// of `heads` is set where a non-empty list starts at position p of the concatenation
// (and at the end of the last list when that is not a multiple of 64), so the owner of
// position p is the number of set bits at or below p, less one: one 64-bit word and a
// lane bit count per window of 64 positions, the windows walked in order with a running
// count of the owners before them.  A window's head bits are also the run starts of
// its segmented scan.  Per owner (its rank among the range's non-empty lists):
// `base`, the list's first edge slot less its position, and a 32-bit value.
// A range holds at most DENSE_RANGE_MAX slots (checked where the tasks are built); the
// heads rows also cover the words read past the end by the windows in flight.
constexpr int RANGE_WORDS = DENSE_RANGE_MAX / WAVE + DENSE_WIN;
struct DenseLds {
    uint64_t heads[BLK / WAVE][RANGE_WORDS];
    int base[BLK / WAVE][DENSE_V];
    int val[BLK / WAVE][DENSE_V];
};

// Builds the wave's range index over its lists (lane l's list j is list j * WAVE + l:
// first edge slot beg[j], length len[j], value val[j]); rank[j] receives the list's
// owner index (meaningful where len[j] > 0).  Returns the total length.
__device__ __forceinline__ int range_index(DenseLds &lds, int wv, int lane, const int (&beg)[DENSE_PER],
                                           const int (&len)[DENSE_PER], const int (&val)[DENSE_PER],
                                           int (&rank)[DENSE_PER]) {
    uint64_t *hd = lds.heads[wv];
    int *bs = lds.base[wv], *vs = lds.val[wv];
    int total = 0, lists = 0, pos[DENSE_PER];
#pragma unroll
    for (int j = 0; j < DENSE_PER; ++j) {
        int incl = len[j];
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
            const int y = __shfl_up(incl, d);
            if (lane >= d) incl += y;
        }
        pos[j] = total + incl - len[j];
        total += __shfl(incl, WAVE - 1);
        const uint64_t ne = wave_ballot(len[j] > 0);
        rank[j] = lists + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(ne >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ne, 0u));
        lists += __popcll(ne);
        if (len[j] > 0) {
            bs[rank[j]] = beg[j] - pos[j];
            vs[rank[j]] = val[j];
        }
    }
    const int words = (total >> 6) + DENSE_WIN;  // <= RANGE_WORDS for total <= DENSE_RANGE_MAX
    for (int i = lane; i < words; i += WAVE) hd[i] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int j = 0; j < DENSE_PER; ++j)
        if (len[j] > 0) atomicOr((unsigned long long *)&hd[pos[j] >> 6], 1ull << (pos[j] & 63));
    if (lane == 0 && (total & 63)) atomicOr((unsigned long long *)&hd[total >> 6], 1ull << (total & 63));
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return total;
}

// The head bits of the window starting at position q (a multiple of 64), as a
// wave-uniform value.
__device__ __forceinline__ uint64_t range_word(const uint64_t *hd, int q) {
    const uint64_t w = hd[q >> 6];
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)w);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(w >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Owner index of this lane's position in a window with head bits w, `before` owners
// having started before the window.
__device__ __forceinline__ int range_owner(uint64_t w, int before, int lane) {
    const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(w >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)w, 0u));
    return before + below + (int)((w >> lane) & 1) - 1;
}

template <bool EID>
__device__ __forceinline__ void dense_range(const SsspArgs &a, int L, int va, int vb, long long fmin, DenseLds &lds,
                                            long long &nf, long long &mf, long long &ex, long long &mn) {
    const FrontierArgs &fa = a.fa;
    uint32_t *nx = fa.bm[(L + 1) % 3];
    const int *adj = fa.adj[2];
    const int *eid = fa.eid[2];
    const int *off = fa.off[2];
    const int lane = lane_id(), wv = (int)threadIdx.x / WAVE;
    const uint64_t *hd = lds.heads[wv];
    const int *bs = lds.base[wv];
    int *mins = lds.val[wv];
    const uint64_t upto = lane == WAVE - 1 ? ~0ull : ((2ull << lane) - 1);  // lanes 0..lane
    const int fmin32 = fmin < INT_MAX ? (int)fmin : INT_MAX;
    // vertex va + 64 j + lane is list 64 j + lane
    int beg[DENSE_PER], len[DENSE_PER], cur[DENSE_PER], rank[DENSE_PER];
#pragma unroll
    for (int j = 0; j < DENSE_PER; ++j) {
        const int v = va + j * WAVE + lane;
        const bool in = v < vb;
        const int vv = in ? v : va;
        const int o0 = off[vv], o1 = off[vv + 1];
        cur[j] = a.dist[vv];
        beg[j] = o0;
        // a vertex at or below fmin + (smallest weight) cannot improve: its list is skipped
        len[j] = in && o1 - o0 <= DENSE_BIG && fmin + a.wmin < cur[j] ? o1 - o0 : 0;
    }
    const int total = range_index(lds, wv, lane, beg, len, cur, rank);
    if (lane == 0) ex += total;  // the wave's edges, counted once
    // an in-bounds slot for the lanes past the end: off[va] (the windows run only when
    // some list of the range is non-empty, and it starts at or after off[va])
    const int safe = __builtin_amdgcn_readfirstlane(beg[0]);
    int before = 0;
    for (int p0 = 0; p0 < total; p0 += DENSE_WIN * WAVE) {
        int p[DENSE_WIN], own[DENSE_WIN], slot[DENSE_WIN], bound[DENSE_WIN];
        uint64_t w[DENSE_WIN];
        bool ok[DENSE_WIN];
#pragma unroll
        for (int k = 0; k < DENSE_WIN; ++k) {
            w[k] = range_word(hd, p0 + k * WAVE);
            p[k] = p0 + k * WAVE + lane;
            ok[k] = p[k] < total;
            own[k] = ok[k] ? range_owner(w[k], before, lane) : 0;
            before += __popcll(w[k]);
            slot[k] = ok[k] ? bs[own[k]] + p[k] : safe;
            bound[k] = mins[own[k]];
        }
        int u[DENSE_WIN], wt[DENSE_WIN];
#pragma unroll
        for (int k = 0; k < DENSE_WIN; ++k) u[k] = adj[slot[k]];
        if (EID) {
            int ei[DENSE_WIN];
#pragma unroll
            for (int k = 0; k < DENSE_WIN; ++k) ei[k] = eid[slot[k]];
#pragma unroll
            for (int k = 0; k < DENSE_WIN; ++k) wt[k] = a.w[ei[k]];
        } else {
#pragma unroll
            for (int k = 0; k < DENSE_WIN; ++k) wt[k] = a.w[slot[k]];
        }
        // 32-bit forms: fmin + w < bound as w < bound - fmin (weights are non-negative,
        // fmin32 in [0, INT_MAX]), and dist[u] + w in unsigned arithmetic (an unreached
        // INT_MAX plus a weight below INT_MAX does not wrap)
        bool need[DENSE_WIN];
        int du[DENSE_WIN];
#pragma unroll
        for (int k = 0; k < DENSE_WIN; ++k) need[k] = ok[k] && wt[k] < bound[k] - fmin32;
#pragma unroll
        for (int k = 0; k < DENSE_WIN; ++k) du[k] = a.dist[need[k] ? u[k] : va];
#pragma unroll
        for (int k = 0; k < DENSE_WIN; ++k) {
            const uint32_t d = (uint32_t)du[k] + (uint32_t)wt[k];
            int best = need[k] && d < (uint32_t)INT_MAX ? (int)d : INT_MAX;
            // lanes with a candidate below their owner's minimum: none (the common case
            // once the bound prunes) skips the window's reduction; a few fold theirs in
            // with LDS atomics; otherwise a segmented scan reduces each owner's run first
            const bool cand = ok[k] && best < bound[k];
            const uint64_t cm = wave_ballot(cand);
            if (cm == 0) continue;
            if (__popcll(cm) <= DENSE_DIRECT) {
                if (cand) atomicMin(&mins[own[k]], best);
                continue;
            }
            // runs: one per owner (lane 0 starts one); the lanes past the end follow the
            // end bit and form one more run
            const int start = 63 - __builtin_clzll((w[k] | 1ull) & upto);
            best = seg_min_scan(best, lane, start);
            const bool tail = lane == WAVE - 1 || ((w[k] >> lane) & 2ull);
            if (ok[k] && tail && best < bound[k]) atomicMin(&mins[own[k]], best);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // write-back: the improved distances, then the degrees of the improved vertices,
    // each stage's loads in flight together (clamped to va where not needed)
    bool found[DENSE_PER];
#pragma unroll
    for (int j = 0; j < DENSE_PER; ++j) {
        const int v = va + j * WAVE + lane;
        const int best = len[j] > 0 ? mins[rank[j]] : INT_MAX;
        found[j] = len[j] > 0 && best < cur[j];
        if (found[j]) {
            a.dist[v] = best;
            mn = best < mn ? best : mn;
        }
    }
    int o0[DENSE_PER], o1[DENSE_PER], p0[DENSE_PER], p1[DENSE_PER];
#pragma unroll
    for (int j = 0; j < DENSE_PER; ++j) {
        const int d = found[j] ? va + j * WAVE + lane : va;
        o0[j] = fa.off[0][d];
        o1[j] = fa.off[0][d + 1];
        if (fa.nviews > 1) {
            p0[j] = fa.off[1][d];
            p1[j] = fa.off[1][d + 1];
        }
    }
#pragma unroll
    for (int j = 0; j < DENSE_PER; ++j) {
        const int v = va + j * WAVE + lane;
        if (found[j]) {
            ++nf;
            mf += o1[j] - o0[j];
            if (fa.nviews > 1) mf += p1[j] - p0[j];
        }
        const uint64_t fm = wave_ballot(found[j]);
        const uint32_t mine = (lane < 32) ? (uint32_t)fm : (uint32_t)(fm >> 32);
        if ((lane & 31) == 0 && mine) atomicOr(&nx[v >> 5], mine);
    }
    // the next task rewrites this wave's LDS rows
    __builtin_amdgcn_wave_barrier();
}

template <bool EID>
__device__ __forceinline__ void dense_pull(const SsspArgs &a, int L, DenseLds &lds, long long &nf, long long &mf,
                                           long long &ex, long long &mn) {
    const int lane = lane_id();
    const int wave = (blockIdx.x * BLK + (int)threadIdx.x) / WAVE, nwaves = gridDim.x * BLK / WAVE;
    const long long fmin = a.fa.ctl->fmin;
    for (int t = wave; t < a.ndpull; t += nwaves) {
        const ChunkRec r = a.dpull[t];
        if (r.v < 0) {
            dense_range<EID>(a, L, r.beg, r.end, fmin, lds, nf, mf, ex, mn);
            continue;
        }
        ex += (r.end - r.beg + WAVE - 1 - lane) / WAVE;
        const long long dv = a.dist[r.v];
        long long best = pull_best<EID>(a, r.beg + lane, r.end, WAVE, fmin, dv, r.v);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const long long y = __shfl_xor(best, o);
            best = y < best ? y : best;
        }
        if (lane == 0 && best < dv && best < a.dist[r.v]) {
            const int old = atomicMin(&a.dist[r.v], (int)best);
            mn = best < mn ? best : mn;
            if (best < old && set_bit_first(a.fa.bm[(L + 1) % 3], r.v)) {
                ++nf;
                mf += view_degree(a.fa, r.v);
            }
        }
    }
}

// Wide push levels of a symmetric graph (frontier out-edges above m / 16) reuse the
// dense tasks, restricted to frontier vertices: a range task concatenates the lists of
// its frontier vertices (the same range index as the dense pull,
// so neighbour ids and weights are coalesced loads), a piece of a frontier hub's list
// is walked by its wave; every edge relaxes its target with atomicMin and marks it.
// One launch balances hubs and low-degree vertices, instead of lanes per vertex.
__device__ __forceinline__ bool range_push_level(const SsspArgs &a, const Ctl *c) {
    return a.push_range && c->mf_cur * 16 > c->m_unexplored;
}

__device__ __forceinline__ void range_push(const SsspArgs &a, int L, DenseLds &lds, long long &nf, long long &mf,
                                           long long &ex, long long &mn) {
    const FrontierArgs &fa = a.fa;
    const uint32_t *fr = fa.bm[L % 3];
    uint32_t *nx = fa.bm[(L + 1) % 3];
    const int *adj = fa.adj[0];
    const int *off = fa.off[0];
    const int lane = lane_id(), wv = (int)threadIdx.x / WAVE;
    const uint64_t *hd = lds.heads[wv];
    const int *bs = lds.base[wv], *dsrc = lds.val[wv];
    const int wave = (blockIdx.x * BLK + (int)threadIdx.x) / WAVE, nwaves = gridDim.x * BLK / WAVE;
    const long long thr = a.bucketed ? fa.ctl->thr : UNBUCKETED;
    // a frontier vertex at or above the bucket bound waits in the next frontier (each
    // piece of a long list checks it; the first to set the bit counts it)
    auto defer = [&](int v, long long dv) {
        mn = dv < mn ? dv : mn;
        if (set_bit_first(nx, v)) {
            ++nf;
            mf += view_degree(fa, v);
        }
    };
    for (int t = wave; t < a.ndpull; t += nwaves) {
        const ChunkRec r = a.dpull[t];
        if (r.v >= 0) {
            if (!test_bit(fr, r.v)) continue;  // wave-uniform
            const long long dv = a.dist[r.v];
            if (dv >= thr) {  // wave-uniform
                if (lane == 0) defer(r.v, dv);
                continue;
            }
            for (int e = r.beg + lane; e < r.end; e += 4 * WAVE) {
                int sl[4], u[4], wt[4], du[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) sl[k] = e + k * WAVE < r.end ? e + k * WAVE : e;
#pragma unroll
                for (int k = 0; k < 4; ++k) u[k] = adj[sl[k]];
#pragma unroll
                for (int k = 0; k < 4; ++k) wt[k] = a.w[sl[k]];
#pragma unroll
                for (int k = 0; k < 4; ++k) du[k] = a.dist[u[k]];
                long long nd[4];
                bool ok[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    ok[k] = e + k * WAVE < r.end;
                    nd[k] = dv + wt[k];
                    ex += ok[k];
                }
                relax_k<4>(a, nx, u, nd, du, ok, nf, mf, mn);
            }
            continue;
        }
        const int va = r.beg, vb = r.end;
        int beg[DENSE_PER], len[DENSE_PER], dvs[DENSE_PER], rank[DENSE_PER];
#pragma unroll
        for (int j = 0; j < DENSE_PER; ++j) {
            const int v = va + j * WAVE + lane;
            const int vv = v < vb ? v : va;
            const int o0 = off[vv], o1 = off[vv + 1];
            beg[j] = o0;
            len[j] = v < vb && o1 - o0 <= DENSE_BIG && test_bit(fr, v) ? o1 - o0 : 0;
            dvs[j] = a.dist[vv];
        }
#pragma unroll
        for (int j = 0; j < DENSE_PER; ++j)
            if (len[j] > 0 && dvs[j] >= thr) {
                defer(va + j * WAVE + lane, dvs[j]);
                len[j] = 0;
            }
        const int total = range_index(lds, wv, lane, beg, len, dvs, rank);
        const int safe = __builtin_amdgcn_readfirstlane(beg[0]);  // off[va]: in bounds when total > 0
        int before = 0;
        for (int p0 = 0; p0 < total; p0 += DENSE_WIN * WAVE) {
            int p[DENSE_WIN], own[DENSE_WIN], sl[DENSE_WIN], u[DENSE_WIN], wt[DENSE_WIN], du[DENSE_WIN];
            bool ok[DENSE_WIN];
#pragma unroll
            for (int k = 0; k < DENSE_WIN; ++k) {
                const uint64_t w = range_word(hd, p0 + k * WAVE);
                p[k] = p0 + k * WAVE + lane;
                ok[k] = p[k] < total;
                own[k] = ok[k] ? range_owner(w, before, lane) : 0;
                before += __popcll(w);
                sl[k] = ok[k] ? bs[own[k]] + p[k] : safe;
            }
#pragma unroll
            for (int k = 0; k < DENSE_WIN; ++k) u[k] = adj[sl[k]];
#pragma unroll
            for (int k = 0; k < DENSE_WIN; ++k) wt[k] = a.w[sl[k]];
#pragma unroll
            for (int k = 0; k < DENSE_WIN; ++k) du[k] = a.dist[u[k]];
            long long nd[DENSE_WIN];
#pragma unroll
            for (int k = 0; k < DENSE_WIN; ++k) {
                nd[k] = (long long)dsrc[own[k]] + wt[k];
                ex += ok[k];
            }
            relax_k<DENSE_WIN>(a, nx, u, nd, du, ok, nf, mf, mn);
        }
        __builtin_amdgcn_wave_barrier();  // the next task rewrites this wave's LDS rows
    }
}

// Closes a whole-GPU level.  nf / mf: the next frontier's size and out-edges
// (waiting vertices included); mn: its smallest distance.  Bucketed: the bound is
// lifted once the pending vertices have fewer than bucket_edges out-edges (the tail of
// low-degree vertices is cheaper to relax unbucketed than at one level per bucket,
// each carrying the waiting vertices along); otherwise, when no pending vertex lies
// below the bound, it moves to the bucket of the smallest one.
__device__ void sssp_close(const SsspArgs &a, int L, long long nf, long long mf, long long mn) {
    Ctl *c = a.fa.ctl;
    int dir = c->dir;
    trace_level(a.fa, L, dir, nf, mf);
    c->fmin = mn;  // every vertex of the next frontier improved (or waits) at mn or above
    if (a.bucketed && c->thr < UNBUCKETED) {
        if (mf < a.bucket_edges) c->thr = UNBUCKETED;
        else if (mn >= c->thr) c->thr = (mn / a.delta + 1) * a.delta;
    }
    if (dir == DIR_PULL) c->bu_levels++;
    if (c->can_bu) {
        const long long m = c->m_unexplored;  // total edges (set at seed)
        if (dir == DIR_PUSH && mf * c->alpha > m) dir = DIR_PULL;
        else if (dir == DIR_PULL && mf * c->beta < m) dir = DIR_PUSH;
    }
    c->dir = dir;
    advance_level(c, L, nf, mf);
}

// Starts a search in one launch: distances (0 at src, INT_MAX elsewhere), the frontier
// bitmaps (src marked in frontier 0) and the far bitmap, the level counters, and the
// control block (workgroup 0).  Every word is written once, so no ordering between
// workgroups is needed.
__global__ void sssp_begin(SsspArgs a, int src, int iter_limit, int can_pull, long long m) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
    for (int v = tid; v < a.fa.n; v += nt) a.dist[v] = v == src ? 0 : INT_MAX;
    for (int i = tid; i < a.fa.nwords; i += nt) {
        a.fa.bm[0][i] = i == (src >> 5) ? 1u << (src & 31) : 0u;
        a.fa.bm[1][i] = 0;
        a.fa.bm[2][i] = 0;
        a.farbm[i] = 0;
    }
    if (blockIdx.x != 0) return;
    clear_parts(a.fa);
    if (threadIdx.x != 0) return;
    Ctl *c = a.fa.ctl;
    a.fa.q[0][0] = src;
    ctl_reset(c, a.fa.n, iter_limit, 1);
    c->thr = a.delta;  // the first bucket: [0, delta)
    c->fmin = 0;
    c->mf_cur = view_degree(a.fa, src);
    // pull only pays on the densest levels: every in-edge is read and the frontier
    // ones gather dist[u]; push gathers only along frontier out-edges
    c->can_bu = can_pull;
    c->alpha = 2;
    c->beta = 2;
    c->m_unexplored = m;
}

// Closes a pending multi-workgroup level, then runs small push levels back to back.
//
// Narrow levels are latency-bound (a grid has thousands of them), so the loop keeps
// the dependent chain per level short.  The frontier (at most SMALL_Q vertices and
// SSSP_SMALL_E edges), its distances, list starts and the prefix sums of its list lengths
// live in LDS; the threads walk the concatenated lists, SMALL_EW edges each in flight,
// finding each slot's owner by binary search over the prefix sums, so a level is one
// round of neighbour/weight loads and one of distance loads whatever the degrees.  The
// next frontier is de-duplicated in an LDS hash set and queued in LDS; the frontier
// bitmaps and the control block are written only when the loop hands over.  Only this
// one workgroup touches the distances while it runs, so loads and atomics take
// workgroup scope (the CU's caches) rather than the device-wide path.
//
// Buckets (delta-stepping, Meyer and Sanders): inside the loop only vertices below the
// threshold thr are relaxed; a vertex improved to thr or more waits in the far list
// (ws.hard, deduplicated by the far bitmap ws.visited).  When no vertex below thr is
// left, thr moves to the bucket of the smallest waiting distance and those below it
// become the frontier.  Every improved vertex is relaxed again later, so the fixed
// point is Bellman-Ford's; on a weighted grid this cuts the relaxations from ~75 per
// edge to ~1.3.  When the loop hands over to the whole-GPU levels (which relax every
// frontier vertex, unbucketed), the queue and the far list become the next bitmap.
constexpr int SMALL_HASH = 2 * SMALL_Q;  // LDS hash set of a level's next frontier
constexpr int SMALL_EW = 4;              // edges per thread in flight
// Largest frontier (out-edges) the loop keeps: one workgroup relaxes ~0.4 G edges/s,
// so wider levels are cheaper as whole-GPU launches (about 25 us for the three).
constexpr int SSSP_SMALL_E = 16384;
constexpr int SMALL_QPT = SMALL_Q / SMALL_BLK;  // queue entries per thread (contiguous)
__global__ __launch_bounds__(SMALL_BLK) void sssp_small(SsspArgs a) {
    Ctl *c = a.fa.ctl;
    if (threadIdx.x == 0 && !c->done) c->steps++;  // read by the host to size the next search's launches
    const FrontierArgs &fa = a.fa;
    __shared__ int s_q[SMALL_Q];             // next-frontier appends (and the entry frontier)
    __shared__ int s_hash[SMALL_HASH];
    __shared__ int s_pre[2 * SMALL_Q + 1];   // list l = view * nq + i: prefix of lengths
    __shared__ int s_beg[2 * SMALL_Q];       // list starts
    __shared__ int s_d[SMALL_Q];             // frontier distances
    __shared__ int s_wave[SMALL_BLK / WAVE];
    __shared__ int s_go, s_L, s_cnt, s_levels, s_limit, s_canpull, s_alpha, s_nfar, s_keep, s_qv, s_nq;
    __shared__ long long s_m, s_thr;
    if (c->pending) {
        const int L = c->level;
        long long nf, mf, ex, mn;
        take_level_parts<SMALL_BLK>(fa, nf, mf, ex, &mn);
        if (threadIdx.x == 0) {
            c->edges_examined += ex;
            sssp_close(a, L, nf, mf, mn);
            c->pending = 0;
            c->qvalid = 0;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int L = c->level;
        s_go = !c->done && c->dir == DIR_PUSH && c->qsize[L % 3] <= SMALL_Q && c->mf_cur <= SSSP_SMALL_E;
        s_L = L;
        s_cnt = c->qsize[L % 3];
        s_levels = 0;
        s_limit = c->iter_limit;
        s_canpull = c->can_bu;
        s_alpha = c->alpha;
        s_m = c->m_unexplored;
        s_thr = c->thr;
        s_nfar = 0;
        s_qv = c->qvalid;
    }
    __syncthreads();
    if (!s_go) return;
    const long long delta = a.delta;
    const int nviews = fa.nviews;
    int *far = a.far;
    uint32_t *farbm = a.farbm;
    long long mf_over = 0;  // out-edges of next-frontier vertices that did not fit the queue
    uint32_t *nx_over = nullptr;  // where such vertices are marked (the next bitmap)
    auto clear_hash = [&]() {
        for (int i = threadIdx.x; i < SMALL_HASH; i += SMALL_BLK) s_hash[i] = -1;
    };
    // u joins the next frontier once: LDS hash set (linear probing, up to 32 slots),
    // then the queue.  A crowded set (only when the next frontier outgrows the queue,
    // which ends the loop) falls back to the next bitmap; past the queue's capacity a
    // vertex is marked in the next bitmap instead of queued.  A vertex let through
    // twice is only relaxed twice.
    // True for the one call that adds u to the next frontier's hash set (or, when the
    // set is crowded, to the next bitmap).
    auto claim = [&](int u) -> bool {
        uint32_t h = ((uint32_t)u * 2654435761u) >> 19;  // 13 bits: SMALL_HASH slots
        for (int probe = 0; probe < 32; ++probe, h = (h + 1) & (SMALL_HASH - 1)) {
            const int old = atomicCAS(&s_hash[h], -1, u);
            if (old == u) return false;  // already in
            if (old == -1) return true;
        }
        return set_bit_first_wg(nx_over, u);
    };
    // Queues claimed vertices (every lane of the wave calls it; `want` selects), one
    // LDS counter reservation per wave; past the queue's capacity a vertex is marked
    // in the next bitmap instead.
    auto enqueue = [&](bool want, int u) {
        const uint64_t mask = wave_ballot(want);
        if (!mask) return;
        const int leader = __ffsll((unsigned long long)mask) - 1;
        int base = 0;
        if (lane_id() == leader) base = atomicAdd(&s_cnt, __popcll(mask));
        base = __shfl(base, leader);
        if (!want) return;
        const int idx = base + lanes_below(mask);
        if (idx < SMALL_Q) {
            s_q[idx] = u;
        } else {
            atomicOr(&nx_over[u >> 5], 1u << (u & 31));
            mf_over += view_degree(fa, u);
        }
    };
    auto append = [&](int u) {
        if (!claim(u)) return;
        const int idx = atomicAdd(&s_cnt, 1);
        if (idx < SMALL_Q) {
            s_q[idx] = u;
        } else {
            atomicOr(&nx_over[u >> 5], 1u << (u & 31));
            mf_over += view_degree(fa, u);
        }
    };
    auto to_far = [&](int u) {
        if (set_bit_first_wg(farbm, u)) far[atomicAdd(&s_nfar, 1)] = u;
    };
    // Moves the threshold to the bucket of the smallest waiting distance and the far
    // vertices below it into the next frontier; far entries below the old threshold
    // were relaxed as frontier vertices already and are dropped.
    auto next_bucket = [&]() {
        const long long old_thr = s_thr;
        const int nfar = s_nfar;
        long long lo = LLONG_MAX;
        for (int i = threadIdx.x; i < nfar; i += SMALL_BLK) {
            const long long d = ld_wg(&a.dist[far[i]]);
            if (d >= old_thr) lo = d < lo ? d : lo;
        }
        lo = block_min<SMALL_BLK>(lo);
        const long long thr = lo == LLONG_MAX ? old_thr : (lo / delta + 1) * delta;
        // stable in-place compaction of the kept entries, SMALL_BLK at a time
        int kept = 0;
        for (int b = 0; b < nfar; b += SMALL_BLK) {
            const int i = b + (int)threadIdx.x;
            const int u = i < nfar ? far[i] : -1;
            bool keep = false;
            if (u >= 0) {
                const long long d = ld_wg(&a.dist[u]);
                keep = d >= thr;
                if (!keep) {
                    clear_bit_wg(farbm, u);
                    if (d >= old_thr) append(u);
                }
            }
            int total;
            const int ex = block_exclusive_scan<SMALL_BLK>(keep ? 1 : 0, s_wave, &total);
            if (keep) far[kept + ex] = u;  // kept + ex <= i: no unread entry is overwritten
            kept += total;
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            s_thr = thr;
            s_nfar = kept;
        }
        __syncthreads();
    };
    // The queued next frontier s_q[0, nq) becomes the frontier: distances, list starts
    // and length prefix sums into LDS.  Returns this thread's share of its out-edges.
    auto build = [&](int nq) -> long long {
        int v[SMALL_QPT], len[2][SMALL_QPT];
        long long mf = 0;
#pragma unroll
        for (int k = 0; k < SMALL_QPT; ++k) {
            const int i = SMALL_QPT * (int)threadIdx.x + k;
            v[k] = i < nq ? s_q[i] : -1;
        }
#pragma unroll
        for (int k = 0; k < SMALL_QPT; ++k) {
            const int i = SMALL_QPT * (int)threadIdx.x + k;
            len[0][k] = len[1][k] = 0;
            if (v[k] < 0) continue;
            s_d[i] = ld_wg(&a.dist[v[k]]);
            for (int view = 0; view < nviews; ++view) {
                const int b = fa.off[view][v[k]];
                s_beg[view * nq + i] = b;
                len[view][k] = fa.off[view][v[k] + 1] - b;
                mf += len[view][k];
            }
        }
        int base = 0;
        for (int view = 0; view < nviews; ++view) {
            int mine = 0;
#pragma unroll
            for (int k = 0; k < SMALL_QPT; ++k) mine += len[view][k];
            int total;
            int ex = block_exclusive_scan<SMALL_BLK>(mine, s_wave, &total) + base;
#pragma unroll
            for (int k = 0; k < SMALL_QPT; ++k) {
                const int i = SMALL_QPT * (int)threadIdx.x + k;
                if (i < nq) s_pre[view * nq + i] = ex;
                ex += len[view][k];
            }
            base += total;
        }
        if (threadIdx.x == 0) {
            s_pre[nviews * nq] = base;
            s_nq = nq;
        }
        return mf;
    };
    {
        const int L = s_L, cur = L % 3, prv = (L + 2) % 3;
        const int qv = s_qv;
        int n0 = s_cnt;
        __syncthreads();
        if (qv & (1 << cur)) {
            for (int i = threadIdx.x; i < n0; i += SMALL_BLK) s_q[i] = fa.q[cur][i];
        } else {
            if (threadIdx.x == 0) s_cnt = 0;
            __syncthreads();
            queue_from_bitmap_to<SMALL_BLK>(fa, L, s_q, &s_cnt, SMALL_Q);
            __syncthreads();
            n0 = min(s_cnt, SMALL_Q);
        }
        if (qv & (1 << prv)) {
            const int *qp = fa.q[prv];
            for (int i = threadIdx.x; i < c->qsize[prv]; i += SMALL_BLK) fa.bm[prv][qp[i] >> 5] = 0;
        } else {
            clear_prev_sweep(fa, L, threadIdx.x, SMALL_BLK);
        }
        clear_hash();
        __syncthreads();
        // the frontier leaves the bitmap for the queue (every bitmap is clean while the
        // loop runs); vertices at or above the threshold wait in the far list
        int vs[SMALL_QPT];
#pragma unroll
        for (int k = 0; k < SMALL_QPT; ++k) {
            const int i = (int)threadIdx.x + k * SMALL_BLK;
            vs[k] = i < n0 ? s_q[i] : -1;
        }
        __syncthreads();
        if (threadIdx.x == 0) s_cnt = 0;
        __syncthreads();
        nx_over = fa.bm[cur];  // cannot overflow: at most n0 <= SMALL_Q entries
#pragma unroll
        for (int k = 0; k < SMALL_QPT; ++k) {
            if (vs[k] < 0) continue;
            fa.bm[cur][vs[k] >> 5] = 0;
            if (ld_wg(&a.dist[vs[k]]) >= s_thr) to_far(vs[k]);
            else append(vs[k]);
        }
        __syncthreads();
        if (s_cnt == 0 && s_nfar > 0) next_bucket();
    }
    build(min(s_cnt, SMALL_Q));
    clear_hash();
    __syncthreads();  // the frontier is in s_d / s_beg / s_pre: the queue is free
    long long ex_all = 0;
    for (;;) {
        const int L = s_L;
        nx_over = fa.bm[(L + 1) % 3];
        const long long thr = s_thr;
        const int nq = s_nq, nl = nviews * nq, T = s_pre[nviews * nq];
        if (threadIdx.x == 0) s_cnt = 0;
        __syncthreads();
        for (int p0 = 0; p0 < T; p0 += SMALL_EW * SMALL_BLK) {
            int p[SMALL_EW], li[SMALL_EW];
#pragma unroll
            for (int k = 0; k < SMALL_EW; ++k) {
                p[k] = p0 + k * SMALL_BLK + (int)threadIdx.x;
                li[k] = 0;
            }
            // owner list = the last one whose prefix is <= p (all lanes' searches in step)
            for (int st = 1 << (31 - __builtin_clz(max(nl, 1))); st > 0; st >>= 1) {
#pragma unroll
                for (int k = 0; k < SMALL_EW; ++k)
                    if (li[k] + st < nl && s_pre[li[k] + st] <= p[k]) li[k] += st;
            }
            int sl[SMALL_EW], u[SMALL_EW], wt[SMALL_EW], du[SMALL_EW], dv[SMALL_EW];
            bool ok[SMALL_EW];
#pragma unroll
            for (int k = 0; k < SMALL_EW; ++k) {
                ok[k] = p[k] < T;
                const int view = li[k] >= nq;  // view-major lists
                sl[k] = ok[k] ? s_beg[li[k]] + (p[k] - s_pre[li[k]]) : 0;  // slot 0 exists: T > 0
                dv[k] = s_d[li[k] - (view ? nq : 0)];
                u[k] = (view ? fa.adj[1] : fa.adj[0])[sl[k]];
            }
#pragma unroll
            for (int k = 0; k < SMALL_EW; ++k) {
                const int *eid = li[k] >= nq ? fa.eid[1] : fa.eid[0];
                wt[k] = a.w[eid ? eid[sl[k]] : sl[k]];
            }
#pragma unroll
            for (int k = 0; k < SMALL_EW; ++k) du[k] = ld_wg(&a.dist[u[k]]);
#pragma unroll
            for (int k = 0; k < SMALL_EW; ++k) {
                bool add = false;
                if (ok[k]) {
                    ++ex_all;
                    const long long nd = (long long)dv[k] + wt[k];
                    if (nd < du[k]) {
                        min_wg(&a.dist[u[k]], (int)nd);
                        if (nd >= thr) to_far(u[k]);
                        else add = claim(u[k]);
                    }
                }
                enqueue(add, u[k]);
            }
        }
        __syncthreads();
        if (s_cnt == 0 && s_nfar > 0) next_bucket();
        const int nf = s_cnt;
        long long mf = mf_over;
        mf_over = 0;
        if (nf <= SMALL_Q) mf += build(nf);  // the next frontier's lists (reads s_q: before any append)
        clear_hash();
        mf = block_sum<SMALL_BLK>(mf);
        if (threadIdx.x == 0) {
            // close level L on the LDS copy of the direction state (as sssp_close)
            trace_level(fa, L, DIR_PUSH, nf, mf);
            s_levels++;
            s_go = nf > 0 && nf <= SMALL_Q && !(s_canpull && mf * s_alpha > s_m) && mf <= SSSP_SMALL_E &&
                   L + 1 < s_limit;
            s_L = L + 1;
            s_keep = nf;
        }
        __syncthreads();
        if (s_go) continue;
        // hand over: the queue and the far list become the next bitmap (the whole-GPU
        // levels relax every frontier vertex), then the control block takes the state
        uint32_t *nx = fa.bm[(L + 1) % 3];
        if (nf <= SMALL_Q) {
            // build() does not run past the queue's capacity; the overflowing vertices'
            // degrees were summed into mf_over as they were marked
            for (int i = threadIdx.x; i < nf; i += SMALL_BLK) atomicOr(&nx[s_q[i] >> 5], 1u << (s_q[i] & 31));
        } else {
            for (int i = threadIdx.x; i < SMALL_Q; i += SMALL_BLK) {
                atomicOr(&nx[s_q[i] >> 5], 1u << (s_q[i] & 31));
            }
        }
        __syncthreads();
        long long nf_far = 0, mf_far = 0, mf_q = 0;
        if (nf > SMALL_Q)  // out-edges of the queued part (build() skipped)
            for (int i = threadIdx.x; i < SMALL_Q; i += SMALL_BLK) mf_q += view_degree(fa, s_q[i]);
        const int nfar = s_nfar;
        for (int i = threadIdx.x; i < nfar; i += SMALL_BLK) {
            const int u = far[i];
            clear_bit_wg(farbm, u);
            if (set_bit_first(nx, u)) {
                ++nf_far;
                mf_far += view_degree(fa, u);
            }
        }
        nf_far = block_sum<SMALL_BLK>(nf_far);
        mf_far = block_sum<SMALL_BLK>(mf_far);
        mf_q = block_sum<SMALL_BLK>(mf_q);
        const long long ex = block_sum<SMALL_BLK>(ex_all);
        if (threadIdx.x == 0) {
            const long long nft = s_keep + nf_far, mft = mf + mf_q + mf_far;
            const int dir = s_canpull && mft * s_alpha > s_m ? DIR_PULL : DIR_PUSH;
            c->edges_examined += ex;
            c->dir = dir;
            c->small_levels += s_levels;
            c->thr = s_thr;
            advance_level(c, L, nft, mft);
            c->qvalid = 0;  // the queue stayed in LDS
        }
        break;
    }
}

__global__ __launch_bounds__(BLK) void sssp_level_a(SsspArgs a, int view) {
    Ctl *c = a.fa.ctl;
    const FrontierArgs &fa = a.fa;
    if (c->done) return;
    const int L = c->level;
    const int dir = c->dir;
    if (dir == DIR_PULL && view != 0) return;
    if (view == 0) clear_prev_sweep(fa, L, blockIdx.x * BLK + threadIdx.x, gridDim.x * BLK);
    if (dir == DIR_PUSH) {
        if (range_push_level(a, c)) {
            return;  // the level is sssp_level_b's
        }
        SsspPush pol(a);
        expand_lanes<BLK>(fa, pol, view, L, fa.bm[L % 3]);
        write_part<BLK>(fa, pol.nf, pol.mf, pol.ex, pol.mn);
        return;
    }
    // a pull level is sssp_level_b's (the dense tasks, one launch)
}

__global__ __launch_bounds__(BLK) void sssp_level_b(SsspArgs a) {
    Ctl *c = a.fa.ctl;
    const FrontierArgs &fa = a.fa;
    if (c->done) return;
    const int L = c->level;
    long long nf = 0, mf = 0, ex = 0, mn = LLONG_MAX;
    __shared__ DenseLds lds;
    if (c->dir == DIR_PUSH && range_push_level(a, c)) {
        range_push(a, L, lds, nf, mf, ex, mn);
    } else if (c->dir == DIR_PUSH) {
        SsspPush pol(a);
        expand_tasks<BLK>(fa, pol, L);
        nf = pol.nf;
        mf = pol.mf;
        ex = pol.ex;
        mn = pol.mn;
    } else if (fa.eid[2]) {
        dense_pull<true>(a, L, lds, nf, mf, ex, mn);
    } else {
        dense_pull<false>(a, L, lds, nf, mf, ex, mn);
    }
    write_part<BLK>(fa, nf, mf, ex, mn);
    if (blockIdx.x == 0 && threadIdx.x == 0) c->pending = 1;
}

}  // namespace mg

using namespace mg;

namespace mg {

// Bucket width of mg_sssp for options o (sssp_delta > 0 as given, < 0 unbucketed,
// 0 automatic: four maximum weights -- a frontier vertex's improvements then stay
// within the next few buckets, and wider buckets mean fewer, fuller rounds).
long long sssp_width(const mg_graph *g, const mg_options &o) {
    if (o.sssp_delta < 0) return UNBUCKETED;
    if (o.sssp_delta > 0) return o.sssp_delta;
    return std::max<long long>(1, 4LL * g->max_weight);
}

// Whole-GPU levels are bucketed only for an explicit width: on R-MAT scale 22
// (weights 1..64) bucket-ordered push levels relax each edge about once (1.0 x m
// against 3 x m) but ran slower than unbucketed levels with dense pulls (2.7 ms at
// width 1, 4.2 ms at width 4, against 2.3 ms), the pushes' atomics and the carried
// waiting vertices costing more than the saved relaxations; the narrow levels are
// always bucketed (grids: 1.12 G relaxations unbucketed, 133 M bucketed).
bool sssp_bucketed_levels(const mg_options &o) { return o.sssp_delta > 0; }

// One SSSP search with the current queues (MG_ERR_OVERFLOW when the edge-chunk queue
// overflowed; the caller regrows it and runs again).
static int sssp_run(mg_graph *g, int src, int *dist_out, const mg_options &o, mg_stats *stats,
                    const CallTimer &timer) {
    int rc;
    Workspace &ws = g->ws;
    SsspArgs a{};
    a.fa = make_frontier_args(g, o.directed);
    a.dist = ws.vals;
    a.w = g->w;
    a.G = o.level_grid > 0 ? o.level_grid : grid_blocks_cap(std::max<int64_t>(g->n / 4, g->m / 64), BLK, BIG_GRID);
    a.delta = sssp_width(g, o);
    a.bucketed = sssp_bucketed_levels(o);
    a.bucket_edges = o.sssp_bucket_edges < 0 ? 0 : o.sssp_bucket_edges > 0 ? (long long)o.sssp_bucket_edges
                                                                            : std::max<long long>(1, g->m / 8);
    a.far = ws.hard;
    a.farbm = ws.visited;
    a.wmin = g->min_weight;
    const int in_view = a.fa.off[2] != nullptr && a.fa.nviews == 1 && o.engine != MG_ENGINE_GAS;
    // pull levels relax every vertex from all its in-edges: unbucketed searches only
    const int can_pull = in_view && !a.bucketed;
    if (in_view) {
        if ((rc = ensure_dense_tasks(g, a.fa.off[2]))) return rc;
        a.dpull = ws.dense_tasks;
        a.ndpull = ws.dense_ntasks;
        // the tasks cut the in-lists; on a symmetric graph those are the out-lists too
        a.push_range = a.fa.off[2] == a.fa.off[0] && a.fa.nviews == 1 && a.fa.eid[0] == nullptr;
    }

    MG_HIP(hipEventRecord(g->ev0, g->stream));
    hipLaunchKernelGGL(sssp_begin, dim3(grid_blocks(g->n, BLK)), dim3(BLK), 0, g->stream, a, src, o.iter_num, can_pull,
                       (long long)g->m);
    MG_HIP(hipGetLastError());
    auto step = [&]() -> int {
        hipLaunchKernelGGL(sssp_small, dim3(1), dim3(SMALL_BLK), 0, g->stream, a);
        for (int v = 0; v < a.fa.nviews; ++v)
            hipLaunchKernelGGL(sssp_level_a, dim3(a.G), dim3(BLK), 0, g->stream, a, v);
        hipLaunchKernelGGL(sssp_level_b, dim3(a.G), dim3(BLK), 0, g->stream, a);
        MG_HIP(hipGetLastError());
        return MG_OK;
    };
    // Unbucketed Bellman-Ford with non-negative weights settles within n rounds and a
    // step runs at least one round; the bucketed narrow loop runs many rounds per step.
    // Twice that bound leaves room; a search that still runs then is reported, not cut.
    const int64_t max_steps = std::min<int64_t>((int64_t)o.iter_num, 2 * (int64_t)g->n + 2) + 3;
    uint64_t key = mix64((uint64_t)o.directed * 4 + (uint64_t)(a.dpull != nullptr) + 2 * (uint64_t)a.bucketed);
    key = mix64(key ^ (uint64_t)a.delta);
    key = mix64(key ^ (uint64_t)a.bucket_edges);
    key = mix64(key ^ (uint64_t)a.G);
    bool tail_queued = false;
    auto tail = [&]() { return queue_finish(g, stats != nullptr, ws.vals, INT_MAX, ws.vals, 4 * (size_t)g->n); };
    if ((rc = drive(g, step, max_steps, KIND_SSSP, key, nullptr, nullptr, tail, &tail_queued)))
        return rc;
    rc = finish_search(g, stats, ws.vals, INT_MAX, dist_out, ws.vals, 4 * (size_t)g->n, o.out_on_device, timer,
                       tail_queued);
    if (rc == MG_OK && o.verbose) rc = print_trace(g, "sssp", "push", "pull");
    return rc;
}

}  // namespace mg

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
    if (o.engine == MG_ENGINE_GAS) return gas_sssp(g, src, dist_out, o, stats, timer);
    g = search_graph(g, o, &src);  // the degree-ordered twin, when the graph keeps one
    if ((rc = ensure_workspace(g))) return rc;
    if ((rc = size_chunk_queue(g, o))) return rc;
    int regrows = 0;
    while ((rc = sssp_run(g, src, dist_out, o, stats, timer)) == MG_ERR_OVERFLOW && regrows < MAX_REGROWS) {
        ++regrows;
        if ((rc = grow_chunk_queue(g))) return rc;
        clear_error();
    }
    if (rc == MG_OK && stats) stats->overflow_regrows = regrows;
    return rc;
}
