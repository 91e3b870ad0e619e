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
// With sorted in-lists the gather is XCD-sliced instead: every in-list is cut into
// 8 ranges by source slice, workgroup b sums the ranges of slice b % 8 (so its
// contribution gathers stay within 1/8 of the array, resident in its XCD's L2),
// and a vertex adds its 8 range sums in slice order.  Ranges up to LANE_T edges
// are summed by one lane in order; longer ones in TASK-edge pieces (lane-strided,
// xor-butterfly) added in piece order.  This order is fixed as well.
#include <vector>

#include "frontier.h"

namespace mg {

constexpr int PR_T1 = 32;
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
    // XCD-sliced gather (sliced = 1): workgroup b gathers from source slice b % 8
    int sliced;
    const int *slice_off;          // (PULL_SLICES - 1) x n cut points in the in-lists
    float *sp;                     // PULL_SLICES x n partial sums, sign bit = an in-neighbour changed
    const ChunkRec *stasks;        // TASK-edge pieces of long ranges, grouped by slice
    const ChunkRec *segs;          // long ranges: {vertex, first piece, pieces, slice}
    float *spart;                  // per piece partial (sign bit as in sp)
    int nsegs;
    int sto[PULL_SLICES + 1];
    // edge-centric gather (edge = 1)
    int edge;
    const int *edge_dst;        // owner of each in-edge slot
    long long m_in;
    const int *span;            // crossing lists: vertex, first chunk, last chunk
    int nspan_short, nspan;
    float *head, *tail;         // per chunk partials (sign bit = an in-neighbour changed)
    int GE, GS2;                // workgroups of pr_e1 and pr_e2 (their partial slots come first)
};

constexpr int PR_SPAN_SHORT = 8;  // crossing lists of up to this many chunks: one thread each

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
        a.contrib[1][v] = a.contrib[0][v];  // vertices without in-edges keep it (edge-centric gather)
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

// One workgroup: closes the iteration from the per-workgroup partials.
__global__ __launch_bounds__(SMALL_BLK) void pr_k3(PrArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const PrIter I = pr_iter(a);
    double delta = 0;
    int nchanged = 0;
    long long nch = nchanged;
    const int np = a.G1 + a.G2 + a.G3;
    for (int i = threadIdx.x; i < np; i += SMALL_BLK) {
        delta += a.partials[i];
        nch += a.cpart[i].nf;
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

// ---- XCD-sliced gather.  Partial sums are non-negative (contributions are ranks
// over positive out-degrees), so the sign bit carries "an in-neighbour changed".
__device__ __forceinline__ float with_act(float s, bool act) {
    return __int_as_float(__float_as_int(s) | (act ? (int)0x80000000 : 0));
}

// Short (vertex, slice) ranges, one lane each in vertex order: workgroup b covers
// slice b % 8, so contribution and frontier-bit gathers stay in one slice (L2).
__global__ __launch_bounds__(BLK) void pr_s1(PrArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const PrIter I = pr_iter(a);
    uint32_t *clr = a.chg[(I.it + 1) % 3];  // the bitmap filled at iteration it+1
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.nwords; i += gridDim.x * blockDim.x) clr[i] = 0;
    const int x = blockIdx.x % PULL_SLICES, nbx = gridDim.x / PULL_SLICES;
    const int *c0 = x == 0 ? nullptr : a.slice_off + (int64_t)(x - 1) * a.n;
    const int *c1 = x == PULL_SLICES - 1 ? nullptr : a.slice_off + (int64_t)x * a.n;
    float *sp = a.sp + (int64_t)x * a.n;
    for (int base = (blockIdx.x / PULL_SLICES) * BLK; base < a.n; base += nbx * BLK) {
        const int v = base + (int)threadIdx.x;
        if (v >= a.n) continue;
        const int beg = c0 ? c0[v] : a.in_off[v], end = c1 ? c1[v] : a.in_off[v + 1];
        if (end - beg > LANE_T) continue;  // pieces: pr_s2
        float s = 0.0f;
        bool act = false;
        for (int e = beg; e < end; ++e) {
            const int u = a.in_adj[e];
            s = s + I.c_old[u];
            act = act || test_bit(I.prev, u);
        }
        sp[v] = with_act(s, act);
    }
}

// Pieces of long ranges, one wave each, on the workgroups of their slice.
__global__ __launch_bounds__(BLK) void pr_s2(PrArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const PrIter I = pr_iter(a);
    const int lane = lane_id();
    const int x = blockIdx.x % PULL_SLICES;
    const int wave = (blockIdx.x / PULL_SLICES) * (BLK / WAVE) + (int)threadIdx.x / WAVE;
    const int nwaves = (gridDim.x / PULL_SLICES) * (BLK / WAVE);
    for (int t = a.sto[x] + wave; t < a.sto[x + 1]; t += nwaves) {
        const ChunkRec r = a.stasks[t];
        bool act = false;
        const float s = wave_gather(a, I, r.beg, r.end, act);
        if (lane == 0) a.spart[t] = with_act(s, act);
    }
}

// Long ranges: pieces added in order.
__global__ __launch_bounds__(BLK) void pr_s3(PrArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    for (int k = blockIdx.x * BLK + (int)threadIdx.x; k < a.nsegs; k += gridDim.x * BLK) {
        const ChunkRec sg = a.segs[k];
        float s = 0.0f;
        bool act = false;
        for (int t = sg.beg; t < sg.beg + sg.end; ++t) {
            const float p = a.spart[t];
            s = s + fabsf(p);
            act = act || signbit(p);
        }
        a.sp[(int64_t)sg.pad * a.n + sg.v] = with_act(s, act);
    }
}

// Every vertex: slices added in order, then the apply.
__global__ __launch_bounds__(BLK) void pr_s4(PrArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const PrIter I = pr_iter(a);
    const bool all = a.mode == MG_PR_ALL || I.it == 0;
    double delta = 0;
    int nchanged = 0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < a.n; v += gridDim.x * blockDim.x) {
        float s = 0.0f;
        bool act = all;
#pragma unroll
        for (int x = 0; x < PULL_SLICES; ++x) {
            const float p = a.sp[(int64_t)x * a.n + v];
            s = s + fabsf(p);
            act = act || signbit(p);
        }
        delta += pr_apply(a, I, v, s, act, nchanged);
    }
    pr_write_part<BLK>(a, blockIdx.x, delta, nchanged);
}

// ---- edge-centric gather.  One wave per PULL_CHUNK consecutive in-edge slots, four
// per lane with all loads in flight.  Each run of slots owned by one vertex is summed
// in slot order by its first lane; a vertex whose whole in-list lies in the chunk is
// applied at once, the partial of a list crossing a chunk boundary goes to the
// chunk's head (first run) or tail (last run) slot, and pr_e2 / pr_e3 add a crossing
// list's partials in chunk order.  Vertices without in-edges keep the values pr_init
// gave them (their sum is 0 every iteration).
__global__ __launch_bounds__(BLK) void pr_e1(PrArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const PrIter I = pr_iter(a);
    const bool all = a.mode == MG_PR_ALL || I.it == 0;
    uint32_t *clr = a.chg[(I.it + 1) % 3];  // the bitmap filled at iteration it+1
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.nwords; i += gridDim.x * blockDim.x) clr[i] = 0;
    __shared__ int s_d[BLK / WAVE][PULL_CHUNK];
    __shared__ float s_c[BLK / WAVE][PULL_CHUNK];
    const int lane = lane_id(), wv = (int)threadIdx.x / WAVE;
    const long long nch = (a.m_in + PULL_CHUNK - 1) / PULL_CHUNK;
    const int wave = (blockIdx.x * BLK + (int)threadIdx.x) / WAVE, nwaves = gridDim.x * BLK / WAVE;
    double delta = 0;
    int nchanged = 0;
    for (long long ch = wave; ch < nch; ch += nwaves) {
        const long long base = ch * PULL_CHUNK, e0 = base + 4 * lane;
        int u[4], d[4];
        if (e0 + 3 < a.m_in) {
            const int4 u4 = *reinterpret_cast<const int4 *>(a.in_adj + e0);
            const int4 d4 = *reinterpret_cast<const int4 *>(a.edge_dst + e0);
            u[0] = u4.x; u[1] = u4.y; u[2] = u4.z; u[3] = u4.w;
            d[0] = d4.x; d[1] = d4.y; d[2] = d4.z; d[3] = d4.w;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                u[j] = e0 + j < a.m_in ? a.in_adj[e0 + j] : -1;
                d[j] = e0 + j < a.m_in ? a.edge_dst[e0 + j] : -1;
            }
        }
        float cv[4];
        bool ac[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cv[j] = u[j] >= 0 ? I.c_old[u[j]] : 0.0f;
            ac[j] = !all && u[j] >= 0 && test_bit(I.prev, u[j]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            s_d[wv][4 * lane + j] = d[j];
            s_c[wv][4 * lane + j] = with_act(cv[j], ac[j]);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int i = 4 * lane + j;
            const int v = d[j];
            if (v < 0 || (i > 0 && s_d[wv][i - 1] == v)) continue;  // not the first slot of its run
            float sum = 0.0f;
            bool act = false;
            int k = i;
            for (; k < PULL_CHUNK && s_d[wv][k] == v; ++k) {
                const float p = s_c[wv][k];
                sum = sum + fabsf(p);
                act = act || signbit(p);
            }
            const bool starts = i > 0 || a.in_off[v] == base;
            const bool ends = k < PULL_CHUNK || a.in_off[v + 1] == base + PULL_CHUNK;
            if (starts && ends) delta += pr_apply(a, I, v, sum, act || all, nchanged);
            else if (!starts) a.head[ch] = with_act(sum, act);
            else a.tail[ch] = with_act(sum, act);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    pr_write_part<BLK>(a, blockIdx.x, delta, nchanged);
}

// In-lists crossing up to PR_SPAN_SHORT chunks: one thread adds the partials in order.
__global__ __launch_bounds__(BLK) void pr_e2(PrArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const PrIter I = pr_iter(a);
    const bool all = a.mode == MG_PR_ALL || I.it == 0;
    double delta = 0;
    int nchanged = 0;
    for (int k = blockIdx.x * BLK + (int)threadIdx.x; k < a.nspan_short; k += gridDim.x * BLK) {
        const int v = a.span[3 * k], cf = a.span[3 * k + 1], cl = a.span[3 * k + 2];
        float p[PR_SPAN_SHORT];
        p[0] = a.tail[cf];
#pragma unroll
        for (int j = 1; j < PR_SPAN_SHORT; ++j) p[j] = cf + j <= cl ? a.head[cf + j] : 0.0f;
        float sum = 0.0f;
        bool act = false;
#pragma unroll
        for (int j = 0; j < PR_SPAN_SHORT; ++j)
            if (cf + j <= cl) {
                sum = sum + fabsf(p[j]);
                act = act || signbit(p[j]);
            }
        delta += pr_apply(a, I, v, sum, act || all, nchanged);
    }
    pr_write_part<BLK>(a, a.GE + blockIdx.x, delta, nchanged);
}

// Longer crossing lists: one wave each, lane-strided over the partials then an
// xor-butterfly (a fixed order).
__global__ __launch_bounds__(BLK) void pr_e3(PrArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const PrIter I = pr_iter(a);
    const bool all = a.mode == MG_PR_ALL || I.it == 0;
    const int lane = lane_id();
    const int wave = (blockIdx.x * BLK + (int)threadIdx.x) / WAVE, nwaves = gridDim.x * BLK / WAVE;
    double delta = 0;
    int nchanged = 0;
    for (int k = a.nspan_short + wave; k < a.nspan; k += nwaves) {
        const int v = a.span[3 * k], cf = a.span[3 * k + 1], cl = a.span[3 * k + 2];
        float sum = 0.0f;
        bool act = false;
        for (int j = lane; j <= cl - cf; j += WAVE) {
            const float p = j == 0 ? a.tail[cf] : a.head[cf + j];
            sum = sum + fabsf(p);
            act = act || signbit(p);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) sum = sum + __shfl_xor(sum, o);
        act = wave_ballot(act) != 0;
        if (lane == 0) delta += pr_apply(a, I, v, sum, act || all, nchanged);
    }
    pr_write_part<BLK>(a, a.GE + a.GS2 + blockIdx.x, delta, nchanged);
}

// In-lists crossing chunk boundaries, short spans first (host-built once per view).
static int pr_spans(mg_graph *g, const int *d_in_off) {
    Workspace &ws = g->ws;
    if (ws.pr_span_view == d_in_off) return MG_OK;
    const int n = g->n;
    std::vector<int32_t> off((size_t)n + 1);
    MG_HIP(hipMemcpy(off.data(), d_in_off, 4 * ((size_t)n + 1), hipMemcpyDeviceToHost));
    std::vector<int32_t> sh, lo;
    for (int v = 0; v < n; ++v) {
        if (off[v + 1] == off[v]) continue;
        const int cf = off[v] / PULL_CHUNK, cl = (off[v + 1] - 1) / PULL_CHUNK;
        if (cf == cl) continue;
        auto &dst = cl - cf + 1 <= PR_SPAN_SHORT ? sh : lo;
        dst.push_back(v);
        dst.push_back(cf);
        dst.push_back(cl);
    }
    const int nshort = (int)(sh.size() / 3);
    sh.insert(sh.end(), lo.begin(), lo.end());
    const int64_t nch = (g->m + PULL_CHUNK - 1) / PULL_CHUNK;
    free_dev(ws.pr_span);
    free_dev(ws.pr_head);
    free_dev(ws.pr_tail);
    ws.pr_span = nullptr;
    ws.pr_head = ws.pr_tail = nullptr;
    ws.pr_span_view = nullptr;
    int rc;
    if ((rc = alloc_i32(&ws.pr_span, sh.size())) || (rc = alloc_f32(&ws.pr_head, (size_t)nch)) ||
        (rc = alloc_f32(&ws.pr_tail, (size_t)nch)))
        return rc;
    if (!sh.empty()) MG_HIP(hipMemcpy(ws.pr_span, sh.data(), 4 * sh.size(), hipMemcpyHostToDevice));
    ws.pr_nspan_short = nshort;
    ws.pr_nspan = (int)(sh.size() / 3);
    ws.pr_span_view = d_in_off;
    return MG_OK;
}

// Builds (once per graph and in-edge view) the medium-vertex list and the chunk
// records of high in-degree vertices, on the host from the in-edge offsets.
static int pr_buckets(mg_graph *g, const int *d_in_off) {
    Workspace &ws = g->ws;
    if (ws.pr_view == d_in_off) return MG_OK;
    const int n = g->n;
    std::vector<int32_t> off((size_t)n + 1);
    MG_HIP(hipMemcpy(off.data(), d_in_off, 4 * ((size_t)n + 1), hipMemcpyDeviceToHost));
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
    if (!mid.empty()) MG_HIP(hipMemcpy(ws.pr_mid, mid.data(), 4 * mid.size(), hipMemcpyHostToDevice));
    if (!ch.empty()) MG_HIP(hipMemcpy(ws.pr_chunks, ch.data(), sizeof(ChunkRec) * ch.size(), hipMemcpyHostToDevice));
    MG_HIP(hipMemcpy(ws.pr_bigv, bigv.data(), 4 * bigv.size(), hipMemcpyHostToDevice));
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
    if ((rc = ensure_pr_workspace(g))) return rc;
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

    int GS = 0, GE3 = 0;  // sliced gather grid (a multiple of PULL_SLICES); pr_e3 grid
    if (edge_pull_enabled()) {
        if ((rc = ensure_edge_dst(g, in_off)) || (rc = pr_spans(g, in_off))) return rc;
        a.edge = 1;
        a.edge_dst = ws.edge_dst;
        a.m_in = g->m;
        a.span = ws.pr_span;
        a.nspan_short = ws.pr_nspan_short;
        a.nspan = ws.pr_nspan;
        a.head = ws.pr_head;
        a.tail = ws.pr_tail;
        a.GE = grid_blocks((g->m + PULL_CHUNK - 1) / PULL_CHUNK * WAVE, BLK);
        a.GS2 = grid_blocks(std::max(a.nspan_short, 1), BLK);
        GE3 = grid_blocks((int64_t)std::max(a.nspan - a.nspan_short, 1) * WAVE, BLK);
        a.G1 = a.GE + a.GS2 + GE3;  // pr_k3 sums the partial slots of all three
        a.G2 = 0;
    } else if (pull_slicing_enabled()) {
        if ((rc = ensure_pull_slices(g, in_off, in_adj))) return rc;
        if (ws.slice_ok) {
            a.sliced = 1;
            a.slice_off = ws.slice_off;
            a.sp = reinterpret_cast<float *>(ws.slice_min);
            a.stasks = ws.slice_tasks;
            a.segs = ws.slice_segs;
            a.spart = ws.slice_part;
            a.nsegs = ws.nslice_segs;
            for (int x = 0; x <= PULL_SLICES; ++x) a.sto[x] = ws.slice_task_off[x];
            GS = std::max(PULL_SLICES, grid_blocks(std::max<int64_t>(g->n, g->m / 16), BLK) / PULL_SLICES * PULL_SLICES);
            a.G2 = 0;  // pr_s4's partials use slots [0, G1)
        }
    }

    // high in-degree vertices' chunk sums (the default gather only)
    a.G3 = !a.edge && !a.sliced && a.nbig > 0 ? grid_blocks(a.nbig, BLK) : 0;
    if (!a.edge && !a.sliced) {
        // The default gather is latency-bound: with a grid-stride loop each thread walks
        // several vertices (pr_k1) or each wave several lists (pr_k2) one after another.
        // Up to PR_MAX_GRID workgroups give every vertex / list its own thread / wave and
        // let the hardware overlap them; the partial slots grow to match.
        auto blocks = [](int64_t items) { return (int)std::max<int64_t>(1, std::min<int64_t>((items + BLK - 1) / BLK, PR_MAX_GRID)); };
        a.G1 = blocks(g->n);
        a.G2 = blocks((int64_t)(a.nmid + a.nchunks) * WAVE);
        const int np = a.G1 + a.G2 + a.G3;
        if (ws.pr_nslots < np) {
            free_dev(ws.pr_partials);
            free_dev(ws.pr_cpart);
            ws.pr_partials = nullptr;
            ws.pr_cpart = nullptr;
            ws.pr_nslots = 0;
            if ((rc = alloc_f64(&ws.pr_partials, np)) || (rc = alloc_parts(&ws.pr_cpart, np))) return rc;
            ws.pr_nslots = np;
        }
        a.partials = ws.pr_partials;
        a.cpart = ws.pr_cpart;
    }

    MG_HIP(hipEventRecord(g->ev0, g->stream));
    hipLaunchKernelGGL(pr_init, dim3(grid_blocks(g->n, BLK)), dim3(BLK), 0, g->stream, a, o.iter_num);
    MG_HIP(hipGetLastError());
    auto step = [&]() -> int {
        if (a.edge) {
            hipLaunchKernelGGL(pr_e1, dim3(a.GE), dim3(BLK), 0, g->stream, a);
            hipLaunchKernelGGL(pr_e2, dim3(a.GS2), dim3(BLK), 0, g->stream, a);
            hipLaunchKernelGGL(pr_e3, dim3(GE3), dim3(BLK), 0, g->stream, a);
        } else if (a.sliced) {
            hipLaunchKernelGGL(pr_s1, dim3(GS), dim3(BLK), 0, g->stream, a);
            hipLaunchKernelGGL(pr_s2, dim3(GS), dim3(BLK), 0, g->stream, a);
            if (a.nsegs) hipLaunchKernelGGL(pr_s3, dim3(grid_blocks(a.nsegs, BLK)), dim3(BLK), 0, g->stream, a);
            hipLaunchKernelGGL(pr_s4, dim3(a.G1), dim3(BLK), 0, g->stream, a);
        } else {
            hipLaunchKernelGGL(pr_k1, dim3(a.G1), dim3(BLK), 0, g->stream, a);
            hipLaunchKernelGGL(pr_k2, dim3(a.G2), dim3(BLK), 0, g->stream, a);
        }
        if (a.G3) hipLaunchKernelGGL(pr_k3b, dim3(a.G3), dim3(BLK), 0, g->stream, a);
        hipLaunchKernelGGL(pr_k3, dim3(1), dim3(SMALL_BLK), 0, g->stream, a);
        MG_HIP(hipGetLastError());
        return MG_OK;
    };
    // the captured steps bake in the per-call constants: key on all of them
    uint64_t key = (uint64_t)a.mode;
    for (uint64_t x : {(uint64_t)__builtin_bit_cast(uint32_t, a.base), (uint64_t)__builtin_bit_cast(uint32_t, a.eps),
                       (uint64_t)__builtin_bit_cast(uint64_t, a.tol_n), (uint64_t)(uintptr_t)in_off,
                       (uint64_t)a.sliced, (uint64_t)a.edge})
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
