// Batched multi-source SSSP: up to 256 sources (slots) relax together over shared
// edge reads.
//
// The reference runs one search per source (Algorithms/SSSP/sssp.cu:310-428).  Here the
// distances are vertex-major, D[v * KS + s] (KS = slots rounded up to 16, one 64-byte
// line per 16 slots), so one wave relaxes one edge for all slots at once: lane l
// owns slots 4l..4l+3 and moves 16 contiguous bytes of D[u].  Each vertex carries
// 256-bit masks F (pairs (v, s) to expand this iteration), N (pairs improved this
// iteration) and FAR (pairs waiting in a later bucket).
//
// Buckets (delta-stepping, width delta): slot s compares the virtual distance
// D + off[s] (off = the bucket base when the slot was seeded) with the bound
// T = (b + 1) * delta of the current bucket b, so every slot runs its own bucket
// sequence on a shared clock.  One iteration:
//   * tasks: every active vertex (bitmap A[L%2]) is cut into SB_TASK-edge tasks;
//   * relax: one wave per task; for each edge (v, u, w) and each s in F[v]:
//     c = D[v][s] + w; if c < D[u][s]: atomicMin(D[u][s], c), then if c + off[s] < T
//     the pair joins N[u] (expanded next iteration) else FAR[u], with u marked in
//     the ring bitmap of bucket (c + off[s]) / delta;
//   * finalize: F = 0 on the old frontier, F[u] = N[u] on the new one.
// When an iteration leaves no pair in N, the next bucket's ring bitmap is scanned
// and its waiting pairs move from FAR to F.  Every improvement is at most
// max_weight above T, so a ring of SB_RING bitmaps never wraps onto a live bucket
// when delta >= max_weight / (SB_RING - 2).  Any such schedule is label-correcting:
// every slot converges to its unique shortest distances, so row s equals mg_sssp
// from the slot's source whatever runs beside it.  A distance read while another
// wave lowers it is only stale-high: the later drop marks the pair again.  Rows
// are transposed out of D when a slot finishes (no frontier pair and no bucket
// above the current one holding its pairs).
#include "batch.h"

#include <climits>
#include <cmath>

namespace mg {

struct SBatchArgs {
    Ctl *ctl;
    SlotCtl *sc;
    const int *off, *adj, *w;
    int *D;                // n x KS
    uint4 *F, *N, *FAR;    // pair masks, 2 uint4 per vertex
    uint32_t *act[2];
    uint32_t *ring;        // SB_RING bitmaps of nwords words
    ChunkRec *tasks;
    int64_t task_cap;
    LevelPart *part;       // [0, G): finalize counts, [G, 2G): relax counts
    uint4 *slot_or;        // per-workgroup OR of new frontier masks
    int *out;              // S x n rows
    int2 *pairs;           // (vertex, slot) pairs of sparse vertices, counted in ctl->nhard
    int64_t pair_cap;
    int n, nwords, KS, G;
    long long delta;       // bucket width (LLONG_MAX / 4: one bucket, plain Bellman-Ford)
};

__device__ __forceinline__ M8 zero8s() { return M8{make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)}; }

__global__ void sb_init(SBatchArgs a, int iter_limit) {
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, nt = (int64_t)gridDim.x * blockDim.x;
    int4 *D4 = reinterpret_cast<int4 *>(a.D);
    const int64_t nd4 = (int64_t)a.n * a.KS / 4;
    for (int64_t i = tid; i < nd4; i += nt) D4[i] = make_int4(INT_MAX, INT_MAX, INT_MAX, INT_MAX);
    for (int64_t i = tid; i < 2 * (int64_t)a.n; i += nt) {
        a.F[i] = make_uint4(0, 0, 0, 0);
        a.N[i] = make_uint4(0, 0, 0, 0);
        a.FAR[i] = make_uint4(0, 0, 0, 0);
    }
    for (int64_t i = tid; i < a.nwords; i += nt) {
        a.act[0][i] = 0;
        a.act[1][i] = 0;
    }
    for (int64_t i = tid; i < (int64_t)SB_RING * a.nwords; i += nt) a.ring[i] = 0;
    int *sci = reinterpret_cast<int *>(a.sc);
    for (int64_t i = tid; i < (int64_t)(sizeof(SlotCtl) / 4); i += nt) sci[i] = 0;
    for (int64_t i = tid; i < 2 * MAX_GRID; i += nt) a.slot_or[i] = make_uint4(0, 0, 0, 0);
    if (tid == 0) {
        ctl_reset(a.ctl, a.n, iter_limit, 0);  // idle until a slot is seeded
        a.ctl->thr = a.delta;
        a.ctl->far_hi = 0;
        a.ctl->pair_end = INT_MAX;
    }
}

// Resets the distance columns (and stray waiting pairs) of the slots about to be seeded.
__global__ void sb_prepare(SBatchArgs a, M8 newmask, SlotList sl) {
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, nt = (int64_t)gridDim.x * blockDim.x;
    for (int64_t u = tid; u < a.n; u += nt) {
        for (int j = 0; j < sl.n; ++j) a.D[u * a.KS + sl.slot[j]] = INT_MAX;
        const M8 f = ld8(a.FAR, (int)u);
        const M8 kept = andnot8(f, newmask);
        if (any8(andnot8(f, kept))) st8(a.FAR, (int)u, kept);
    }
}

// One thread per (slot, source) pair; see batch_seed for the level bookkeeping.
// `defer` (serving): while the search runs, a new slot starts at the next bucket
// instead of the current one -- its source pair waits in FAR (and the ring) and is
// activated when the clock reaches that bucket.  Seeded into the current bucket, a
// slot whose whole search fits in one bucket width keeps that bucket busy, and under
// a steady stream of admissions the clock never advances: older slots waiting in
// later buckets would starve.  Deferred, every slot waits at most for the current
// bucket to drain.
__global__ void sb_seed(SBatchArgs a, SeedList sl, int defer) {
    Ctl *c = a.ctl;
    const int i = threadIdx.x;
    const int pending = c->pending;
    const int Lf = pending ? c->level + 1 : c->level;
    const bool later = defer && !c->done;
    const long long T = c->thr;               // the current bucket's upper bound, the next one's base
    const long long base = later ? T : T - a.delta;  // the slot's bucket base
    if (i < sl.n) {
        const int s = sl.slot[i], v = sl.src[i];
        const uint32_t sb = 1u << (s & 31);
        a.D[(int64_t)v * a.KS + s] = 0;
        if (later) {
            atomicOr(reinterpret_cast<uint32_t *>(a.FAR + 2 * v) + (s >> 5), sb);
            atomicOr(&a.ring[(int64_t)((T / a.delta) % SB_RING) * a.nwords + (v >> 5)], 1u << (v & 31));
        } else {
            atomicOr(reinterpret_cast<uint32_t *>(a.F + 2 * v) + (s >> 5), sb);
            atomicOr(&a.act[Lf & 1][v >> 5], 1u << (v & 31));
        }
        atomicOr(&a.sc->assigned[s >> 5], sb);
        atomicOr(&a.sc->live[s >> 5], sb);
        a.sc->off[s] = base;
        a.sc->far_hi[s] = (unsigned long long)(base / a.delta);
        a.sc->te[s] = 0;
        a.sc->nr[s] = 0;
    }
    __syncthreads();
    if (i == 0 && sl.n > 0) {
        if (later) {
            // the clock must reach the slots' bucket: the close advances while a waiting
            // bucket lies above the current one
            c->far_hi = max(c->far_hi, T / a.delta);
        } else if (pending) {
            atomicAdd((unsigned long long *)&a.part[0].nf, (unsigned long long)sl.n);
        } else {
            c->done = c->iter_limit <= 0;
            c->aux = 0;
        }
    }
}

// One workgroup: closes the previous iteration; moves to the next bucket when the
// frontier ran out (aux = 1 asks sb_activate to refill it).
__global__ __launch_bounds__(SMALL_BLK) void sb_close(SBatchArgs a) {
    Ctl *c = a.ctl;
    if (!c->pending) return;
    long long nf = 0, ex = 0, hi = 0;
    for (int i = threadIdx.x; i < a.G; i += SMALL_BLK) {
        nf += a.part[i].nf;
        const LevelPart p = a.part[a.G + i];
        ex += p.ex;
        hi = max(hi, p.mf);
    }
    nf = block_sum<SMALL_BLK>(nf);
    ex = block_sum<SMALL_BLK>(ex);
    hi = -block_min<SMALL_BLK>(-hi);
    if (threadIdx.x == 0) {
        const int L = c->level;
        c->edges_examined += ex;
        c->far_hi = max(c->far_hi, hi);
        c->pending = 0;
        c->nchunks = 0;
        c->nhard = 0;
        c->pair_end = INT_MAX;
        c->level = L + 1;
        c->nf_prev = (int)nf;
        c->aux = 0;
        if (nf == 0) {
            if (c->thr / a.delta - 1 < c->far_hi) {
                c->thr += a.delta;
                c->aux = 1;
            } else {
                c->done = 1;
            }
        }
        if (L + 1 >= c->iter_limit) c->done = 1;
    }
}

// Calls fn(v, lane) for each set bit of bm, one vertex per wave, and clears the
// words it reads when `clear` is set.
template <class Fn>
__device__ void for_each_bit_wave(uint32_t *bm, int nwords, bool clear, Fn fn) {
    constexpr int WPR = BLK / 2;
    __shared__ int s_q[WPR * 32];
    __shared__ int s_w[BLK / WAVE];
    const int lane = lane_id(), wv = (int)threadIdx.x / WAVE;
    for (int w0 = blockIdx.x * WPR; w0 < nwords; w0 += gridDim.x * WPR) {
        const int w = w0 + (int)threadIdx.x;
        uint32_t word = 0;
        if (threadIdx.x < WPR && w < nwords) {
            word = bm[w];
            if (clear && word) bm[w] = 0;
        }
        int total;
        int k = block_exclusive_scan<BLK>(__popc(word), s_w, &total);
        while (word) {
            s_q[k++] = w * 32 + __ffs(word) - 1;
            word &= word - 1;
        }
        __syncthreads();
        for (int i = wv; i < total; i += BLK / WAVE) fn(s_q[i], lane);
        __syncthreads();
    }
}

// Combines the lanes' slot nibbles into the 32-bit words of a vertex mask.
__device__ __forceinline__ uint32_t nibble_word(uint32_t nib, int lane) {
    uint32_t word = nib << (4 * (lane & 7));
    word |= __shfl_xor(word, 1);
    word |= __shfl_xor(word, 2);
    word |= __shfl_xor(word, 4);
    return word;
}

// ORs the lanes' nibbles into mask m of vertex v (lanes 0, 8, .., 56 issue the atomics).
__device__ __forceinline__ void or_nibbles(uint4 *m, int v, uint32_t nib, int lane) {
    const uint32_t word = nibble_word(nib, lane);
    if ((lane & 7) == 0 && word) atomicOr(reinterpret_cast<uint32_t *>(m + 2 * v) + (lane >> 3), word);
}

__device__ __forceinline__ void mark_active(uint32_t *bm, int v) {
    const uint32_t bit = 1u << (v & 31);
    if (!(bm[v >> 5] & bit)) atomicOr(&bm[v >> 5], bit);
}

// The bucket offsets of lane's four slots.
__device__ __forceinline__ void lane_offsets(const SBatchArgs &a, int lane, long long o[4]) {
#pragma unroll
    for (int b = 0; b < 4; ++b) o[b] = 4 * lane + b < a.KS ? a.sc->off[4 * lane + b] : 0;
}

// New bucket: pairs of the ring bitmap's vertices whose virtual distance lies in
// the bucket move from FAR to F; pairs already below it were expanded when they
// dropped.
__global__ __launch_bounds__(BLK) void sb_activate(SBatchArgs a) {
    Ctl *c = a.ctl;
    if (c->done || !c->aux) return;
    const long long T = c->thr, T0 = T - a.delta;
    const int L = c->level;
    uint32_t *bm = a.ring + (int64_t)((T0 / a.delta) % SB_RING) * a.nwords;
    uint32_t *act = a.act[L & 1];
    long long off4[4];
    lane_offsets(a, lane_id(), off4);
    for_each_bit_wave(bm, a.nwords, true, [&](int u, int lane) {
        const uint32_t fw = reinterpret_cast<const uint32_t *>(a.FAR + 2 * u)[lane >> 3];
        const uint32_t nib = 4 * lane < a.KS ? (fw >> (4 * (lane & 7))) & 0xFu : 0u;
        uint32_t go = 0, drop = 0;
        if (nib) {
            const int4 d = reinterpret_cast<const int4 *>(a.D + (int64_t)u * a.KS)[lane];
            const int dd[4] = {d.x, d.y, d.z, d.w};
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if (nib & (1u << b)) {
                    const long long vd = (long long)dd[b] + off4[b];
                    if (vd < T0) drop |= 1u << b;
                    else if (vd < T) go |= 1u << b;
                }
        }
        const uint32_t gw = nibble_word(go, lane), kw = nibble_word(nib & ~(go | drop), lane);
        if ((lane & 7) == 0 && 4 * lane < a.KS) {
            reinterpret_cast<uint32_t *>(a.FAR + 2 * u)[lane >> 3] = kw;
            if (gw) reinterpret_cast<uint32_t *>(a.F + 2 * u)[lane >> 3] = gw;
        }
        if (__any(go != 0) && lane == 0) atomicOr(&act[u >> 5], 1u << (u & 31));
    });
}

// Active vertices become work: a vertex with few live slots and a short list becomes
// (vertex, slot) pairs, one lane each in sb_relax; others are cut into SB_TASK-edge
// tasks for whole waves (lanes = slots).  Pairs that do not fit the pair buffer
// fall back to tasks.
__global__ __launch_bounds__(BLK) void sb_tasks(SBatchArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const int L = c->level;
    __shared__ int s_w[BLK / WAVE];
    __shared__ int s_tb, s_pb;
    for_each_bit(a.act[L & 1], a.nwords, [&](int v) {
        int beg = 0, deg = 0, pc = 0;
        M8 fv = zero8s();
        if (v >= 0) {
            beg = a.off[v];
            deg = a.off[v + 1] - beg;
            fv = ld8(a.F, v);
            pc = popc8(fv);
        }
        bool as_pairs = v >= 0 && deg > 0 && deg <= LANE_T && pc <= SB_PAIR_MAX;
        int ptot;
        const int pex = block_exclusive_scan<BLK>(as_pairs ? pc : 0, s_w, &ptot);
        if (ptot > 0) {
            if (threadIdx.x == 0) {
                s_pb = atomicAdd(&c->nhard, ptot);
                if ((int64_t)s_pb + ptot > a.pair_cap) atomicMin(&c->pair_end, s_pb);
            }
            __syncthreads();
            const bool fits = (int64_t)s_pb + ptot <= a.pair_cap;
            if (fits && as_pairs) {
                int k = s_pb + pex;
                for (int w = 0; w < 8; ++w) {
                    uint32_t b = word8(fv, w);
                    while (b) {
                        a.pairs[k++] = make_int2(v, w * 32 + __ffs(b) - 1);
                        b &= b - 1;
                    }
                }
            }
            if (!fits) as_pairs = false;
            __syncthreads();
        }
        const int nt = (!as_pairs && deg > 0) ? (deg + SB_TASK - 1) / SB_TASK : 0;
        int tot;
        const int tex = block_exclusive_scan<BLK>(nt, s_w, &tot);
        if (tot > 0) {
            if (threadIdx.x == 0) s_tb = atomicAdd(&c->nchunks, tot);
            __syncthreads();
            for (int t = 0; t < nt; ++t) {
                const int64_t slot = (int64_t)s_tb + tex + t;
                if (slot < a.task_cap)
                    a.tasks[slot] = ChunkRec{v, beg + t * SB_TASK, min(beg + (t + 1) * SB_TASK, beg + deg), 0};
                else
                    atomicOr(&c->changed, 1 << 30);
            }
            __syncthreads();
        }
    });
}

// One wave per task; lanes are slots.
__global__ __launch_bounds__(BLK) void sb_relax(SBatchArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const int L = c->level;
    const long long T = c->thr, delta = a.delta;
    uint32_t *act_next = a.act[(L + 1) & 1];
    const int lane = lane_id();
    const int wave = (blockIdx.x * BLK + (int)threadIdx.x) / WAVE, nwaves = gridDim.x * BLK / WAVE;
    const int nt = (int)min((int64_t)c->nchunks, a.task_cap);
    const bool has_slice = 4 * lane < a.KS;
    long long off4[4], hi4[4] = {-1, -1, -1, -1};
    lane_offsets(a, lane, off4);
    long long ex = 0, hi = 0;
    for (int t = wave; t < nt; t += nwaves) {
        const ChunkRec r = a.tasks[t];
        const uint32_t fw = reinterpret_cast<const uint32_t *>(a.F + 2 * r.v)[lane >> 3];
        const uint32_t nib = has_slice ? (fw >> (4 * (lane & 7))) & 0xFu : 0u;
        if (!__any(nib != 0)) continue;
        int4 dv = make_int4(0, 0, 0, 0);
        if (nib) dv = reinterpret_cast<const int4 *>(a.D + (int64_t)r.v * a.KS)[lane];
        ex += (long long)__popc(nib) * (r.end - r.beg);
        for (int e = r.beg; e < r.end; e += 4) {
            // every load of the group unconditional (clamped to slot e), so they go out together
            int u[4], wt[4];
            int4 du[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int sl = e + j < r.end ? e + j : e;
                u[j] = a.adj[sl];
                wt[j] = a.w[sl];
            }
            if (nib) {
#pragma unroll
                for (int j = 0; j < 4; ++j) du[j] = reinterpret_cast<const int4 *>(a.D + (int64_t)u[j] * a.KS)[lane];
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) du[j] = make_int4(INT_MIN, INT_MIN, INT_MIN, INT_MIN);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (e + j >= r.end) break;  // wave-uniform
                uint32_t near = 0, far = 0, slots = 0;
                if (nib) {
                    const long long cs[4] = {(long long)dv.x + wt[j], (long long)dv.y + wt[j],
                                             (long long)dv.z + wt[j], (long long)dv.w + wt[j]};
                    const int ds[4] = {du[j].x, du[j].y, du[j].z, du[j].w};
                    int *dp = a.D + (int64_t)u[j] * a.KS + 4 * lane;
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        if ((nib & (1u << b)) && cs[b] < ds[b]) {
                            atomicMin(dp + b, (int)cs[b]);
                            const long long vd = cs[b] + off4[b];
                            if (vd < T) {
                                near |= 1u << b;
                            } else {
                                far |= 1u << b;
                                const long long bk = vd / delta;
                                slots |= 1u << (int)(bk % SB_RING);
                                hi = max(hi, bk);
                                hi4[b] = max(hi4[b], bk);
                            }
                        }
                }
                const bool any_near = __any(near != 0), any_far = __any(far != 0);
                if (any_near) {
                    or_nibbles(a.N, u[j], near, lane);
                    if (lane == 0) mark_active(act_next, u[j]);
                }
                if (any_far) {
                    or_nibbles(a.FAR, u[j], far, lane);
#pragma unroll
                    for (int o = 1; o < WAVE; o <<= 1) slots |= __shfl_xor(slots, o);
                    if (lane < SB_RING && (slots >> lane) & 1) mark_active(a.ring + (int64_t)lane * a.nwords, u[j]);
                }
            }
        }
    }
    // pairs: one lane walks a sparse vertex's list for one slot
    const int npairs = (int)min(min((int64_t)c->nhard, (int64_t)c->pair_end), a.pair_cap);
    for (int p = blockIdx.x * BLK + (int)threadIdx.x; p < npairs; p += gridDim.x * BLK) {
        const int2 pr = a.pairs[p];
        const int v = pr.x, sl = pr.y;
        const long long offs = a.sc->off[sl];
        const long long dvs = a.D[(int64_t)v * a.KS + sl];
        const uint32_t bit = 1u << (sl & 31);
        const int beg = a.off[v], end = a.off[v + 1];
        ex += end - beg;
        for (int e = beg; e < end; e += 4) {
            int u[4], wt[4], du[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int s2 = e + j < end ? e + j : e;
                u[j] = a.adj[s2];
                wt[j] = a.w[s2];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) du[j] = a.D[(int64_t)u[j] * a.KS + sl];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const long long cst = dvs + wt[j];
                if (e + j >= end || cst >= du[j]) continue;
                atomicMin(a.D + (int64_t)u[j] * a.KS + sl, (int)cst);
                const long long vd = cst + offs;
                uint32_t *mask = reinterpret_cast<uint32_t *>((vd < T ? a.N : a.FAR) + 2 * u[j]) + (sl >> 5);
                if (!(*mask & bit)) atomicOr(mask, bit);
                if (vd < T) {
                    mark_active(act_next, u[j]);
                } else {
                    const long long bk = vd / delta;
                    mark_active(a.ring + (int64_t)(bk % SB_RING) * a.nwords, u[j]);
                    hi = max(hi, bk);
                    if ((unsigned long long)bk > a.sc->far_hi[sl]) atomicMax(&a.sc->far_hi[sl], (unsigned long long)bk);
                }
            }
        }
    }
    // per-slot highest waiting bucket (read first: most waves add nothing new)
#pragma unroll
    for (int b = 0; b < 4; ++b)
        if (hi4[b] >= 0 && (unsigned long long)hi4[b] > a.sc->far_hi[4 * lane + b])
            atomicMax(&a.sc->far_hi[4 * lane + b], (unsigned long long)hi4[b]);
    ex = block_sum<BLK>(ex);  // each lane counted its own slots
    hi = -block_min<BLK>(-hi);
    if (threadIdx.x == 0) {
        LevelPart p{0, hi, ex, 0};
        a.part[a.G + blockIdx.x] = p;
    }
}

// Finalize: the old frontier's masks are cleared (except for vertices of the new
// frontier, overwritten below) with the old active bitmap, which becomes the next
// level's empty target; the new masks become the frontier.
__global__ __launch_bounds__(BLK) void sb_fin(SBatchArgs a) {
    Ctl *c = a.ctl;
    if (c->done) return;
    const int L = c->level;
    const uint32_t *nxt = a.act[(L + 1) & 1];
    for_each_bit_clear(a.act[L & 1], a.nwords, [&](int v) {
        if (v >= 0 && !test_bit(nxt, v)) st8(a.F, v, zero8s());
    });
    long long nf = 0;
    M8 lv = zero8s();
    for_each_bit(a.act[(L + 1) & 1], a.nwords, [&](int u) {
        if (u < 0) return;
        const M8 nw = ld8(a.N, u);
        st8(a.F, u, nw);
        st8(a.N, u, zero8s());
        lv = or8(lv, nw);
        ++nf;
    });
    nf = block_sum<BLK>(nf);
    // block OR of the new masks
    __shared__ uint4 s_or[2 * (BLK / WAVE)];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lv = or8(lv, shfl_xor8(lv, o));
    if (lane_id() == 0) {
        s_or[2 * (threadIdx.x / WAVE)] = lv.a;
        s_or[2 * (threadIdx.x / WAVE) + 1] = lv.b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        M8 r = zero8s();
        for (int i = 0; i < BLK / WAVE; ++i) r = or8(r, M8{s_or[2 * i], s_or[2 * i + 1]});
        st8(a.slot_or, blockIdx.x, r);
        LevelPart p{nf, 0, 0, 0};
        a.part[blockIdx.x] = p;
        if (blockIdx.x == 0) c->pending = 1;
    }
}

// Distance columns of the listed slots -> their rows, through a 64-vertex x 64-slot
// LDS tile (reads along a vertex's row of D, writes along the slot's row).
__global__ __launch_bounds__(BLK) void sb_extract(SBatchArgs a, SlotList sl) {
    __shared__ int tile[64][65];
    const int tv = (a.n + 63) / 64, ts = (sl.n + 63) / 64;
    for (int t = blockIdx.x; t < tv * ts; t += gridDim.x) {
        const int v0 = (t % tv) * 64, j0 = (t / tv) * 64;
        for (int i = threadIdx.x; i < 64 * 64; i += BLK) {
            const int r = i / 64, cc = i % 64;
            const int v = v0 + r, j = j0 + cc;
            tile[r][cc] = (v < a.n && j < sl.n) ? a.D[(int64_t)v * a.KS + sl.slot[j]] : INT_MAX;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < 64 * 64; i += BLK) {
            const int cc = i / 64, r = i % 64;
            const int v = v0 + r, j = j0 + cc;
            if (v < a.n && j < sl.n) a.out[(int64_t)sl.slot[j] * a.n + v] = tile[r][cc];
        }
        __syncthreads();
    }
}

// Sums the per-slot counts of slots [0, k) into the control block.
__global__ __launch_bounds__(SMALL_BLK) void sb_sum_stats(Ctl *c, const SlotCtl *sc, int k) {
    long long te = 0, nr = 0;
    for (int s = threadIdx.x; s < k; s += SMALL_BLK) {
        te += sc->te[s];
        nr += sc->nr[s];
    }
    te = block_sum<SMALL_BLK>(te);
    nr = block_sum<SMALL_BLK>(nr);
    if (threadIdx.x == 0) {
        c->edges_traversed = te;
        c->reached = nr;
    }
}

static M8 host_mask8(const uint32_t m[8]) {
    return M8{make_uint4(m[0], m[1], m[2], m[3]), make_uint4(m[4], m[5], m[6], m[7])};
}

static SBatchArgs sssp_args(BatchDev &b, mg_graph *g, long long delta) {
    SBatchArgs a{};
    a.ctl = b.ctl;
    a.sc = b.sc;
    a.off = g->row_off;
    a.adj = g->col;
    a.w = g->w;
    a.D = b.D;
    a.F = b.F;
    a.N = b.N;
    a.FAR = b.FAR;
    a.act[0] = b.act[0];
    a.act[1] = b.act[1];
    a.ring = b.ring;
    a.tasks = b.tasks;
    a.task_cap = b.task_cap;
    a.part = b.part;
    a.slot_or = b.slot_or;
    a.out = b.rows;
    a.pairs = b.pairs;
    a.pair_cap = b.pair_cap;
    a.n = g->n;
    a.nwords = b.nwords;
    a.KS = b.KS;
    a.G = b.G;
    a.delta = delta;
    return a;
}

int sssp_batch_reset(BatchDev &b, mg_graph *g, hipStream_t st, int iter_limit, long long delta) {
    SBatchArgs a = sssp_args(b, g, delta);
    hipLaunchKernelGGL(sb_init, dim3(grid_blocks((int64_t)g->n * b.KS / 4 + 2 * (int64_t)g->n, BLK)), dim3(BLK), 0, st, a,
                       iter_limit);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int sssp_batch_seed(BatchDev &b, mg_graph *g, hipStream_t st, const SeedList &sl, long long delta, bool fresh,
                    bool defer) {
    if (sl.n <= 0) return MG_OK;
    SBatchArgs a = sssp_args(b, g, delta);
    if (!fresh) {
        uint32_t m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        SlotList ls{};
        ls.n = sl.n;
        for (int i = 0; i < sl.n; ++i) {
            m[sl.slot[i] >> 5] |= 1u << (sl.slot[i] & 31);
            ls.slot[i] = sl.slot[i];
        }
        hipLaunchKernelGGL(sb_prepare, dim3(grid_blocks(g->n, BLK)), dim3(BLK), 0, st, a, host_mask8(m), ls);
    }
    hipLaunchKernelGGL(sb_seed, dim3(1), dim3(MG_BATCH_MAX), 0, st, a, sl, defer ? 1 : 0);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int sssp_batch_step(BatchDev &b, mg_graph *g, hipStream_t st, long long delta) {
    SBatchArgs a = sssp_args(b, g, delta);
    hipLaunchKernelGGL(sb_close, dim3(1), dim3(SMALL_BLK), 0, st, a);
    hipLaunchKernelGGL(sb_activate, dim3(a.G), dim3(BLK), 0, st, a);
    hipLaunchKernelGGL(sb_tasks, dim3(a.G), dim3(BLK), 0, st, a);
    hipLaunchKernelGGL(sb_relax, dim3(a.G), dim3(BLK), 0, st, a);
    hipLaunchKernelGGL(sb_fin, dim3(a.G), dim3(BLK), 0, st, a);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

int sssp_batch_extract(BatchDev &b, mg_graph *g, hipStream_t st, const SlotList &sl) {
    if (sl.n <= 0) return MG_OK;
    SBatchArgs a = sssp_args(b, g, 1);
    const int tb = grid_blocks((int64_t)((g->n + 63) / 64) * ((sl.n + 63) / 64) * BLK, BLK);
    hipLaunchKernelGGL(sb_extract, dim3(tb), dim3(BLK), 0, st, a, sl);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

// Bucket width: the option, or about four mean edge weights, raised to
// max_weight / (SB_RING - 2); LLONG_MAX / 4 when unbucketed (sssp_delta < 0), which
// no distance reaches.
long long sssp_delta(const mg_graph *g, const mg_options &o) {
    if (o.sssp_delta < 0) return LLONG_MAX / 4;
    const long long d = o.sssp_delta > 0 ? o.sssp_delta : std::llround(4.0 * g->mean_weight);
    // the bucket ring must span every improvement: delta >= max_weight / (SB_RING - 2)
    const long long lo = ((long long)g->max_weight + SB_RING - 3) / (SB_RING - 2);
    return std::max<long long>({1, d, lo});
}

}  // namespace mg

using namespace mg;

namespace mg {

static int sssp_batch_checks(const char *fn, const mg_graph *g, const mg_options &o) {
    if (!g->w) return set_error(MG_ERR_UNSUPPORTED, "%s: the graph has no edge weights", fn);
    if (g->min_weight < 0) return set_error(MG_ERR_UNSUPPORTED, "%s: negative edge weights", fn);
    if (o.directed == 0 && !(g->flags & MG_GRAPH_SYMMETRIC))
        return set_error(MG_ERR_UNSUPPORTED, "%s: directed=0 needs a symmetric graph", fn);
    return MG_OK;
}

// Batched SSSP with npairs (row, source) pairs; every row of [0, k) has a source.
static int sssp_batch_run(mg_graph *g, const int32_t *slot_of, const int32_t *srcs, int32_t npairs, int32_t k,
                          int32_t *dist_out, const mg_options &o, mg_stats *stats, const CallTimer &timer) {
    int rc;
    BatchDev *b = nullptr;
    if ((rc = graph_batch_dev(g, true, &b))) return rc;
    const long long delta = sssp_delta(g, o);
    SlotList all{};
    all.n = k;
    for (int s = 0; s < k; ++s) all.slot[s] = s;
    auto run = [&]() -> int {
        int rc;
        MG_HIP(hipEventRecord(g->ev0, g->stream));
        if ((rc = sssp_batch_reset(*b, g, g->stream, o.iter_num, delta))) return rc;
        // a fresh context: the distance columns are clear, every pair only sets its source
        for (int i0 = 0; i0 < npairs; i0 += MG_BATCH_MAX) {
            SeedList sl{};
            for (int i = i0; i < npairs && i < i0 + MG_BATCH_MAX; ++i) {
                sl.slot[sl.n] = slot_of[i];
                sl.src[sl.n++] = srcs[i];
            }
            if ((rc = sssp_batch_seed(*b, g, g->stream, sl, delta, true, false))) return rc;
        }
        auto step = [&]() -> int { return sssp_batch_step(*b, g, g->stream, delta); };
        if ((rc = drive(g, step, (int64_t)o.iter_num + 3, KIND_BATCH_SSSP, mix64((uint64_t)delta), b->ctl, b->h_ctl)))
            return rc;
        if ((rc = sssp_batch_extract(*b, g, g->stream, all))) return rc;
        if ((rc = batch_row_stats(*b, g, g->stream, all, INT_MAX))) return rc;
        hipLaunchKernelGGL(sb_sum_stats, dim3(1), dim3(SMALL_BLK), 0, g->stream, b->ctl, (const SlotCtl *)b->sc, k);
        MG_HIP(hipGetLastError());
        return finish_batch(g, *b, stats, dist_out, 4 * (size_t)k * g->n, o.out_on_device, timer);
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

extern "C" int mg_sssp_batch(mg_graph *g, const int32_t *srcs, int32_t k, int32_t *dist_out, const mg_options *opt,
                             mg_stats *stats) {
    CallTimer timer;
    mg_options o;
    int rc = begin_search(g, opt, &o);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g->mu);
    if ((rc = check_batch_sources("mg_sssp_batch", g, srcs, k))) return rc;
    if ((rc = sssp_batch_checks("mg_sssp_batch", g, o))) return rc;
    int32_t rows[MG_BATCH_MAX], tsrc[MG_BATCH_MAX];
    for (int s = 0; s < k; ++s) rows[s] = s;
    mg_graph *sg = batch_graph(g, o, srcs, k, tsrc);  // the degree-ordered twin, when the graph keeps one
    if (sg != g && (rc = ensure_workspace(sg))) return rc;
    return sssp_batch_run(sg, rows, sg != g ? tsrc : srcs, k, k, dist_out, o, stats, timer);
}

extern "C" int mg_sssp_batch_sets(mg_graph *g, const int32_t *row_of, const int32_t *srcs, int32_t npairs, int32_t k,
                                  int32_t *dist_out, const mg_options *opt, mg_stats *stats) {
    CallTimer timer;
    mg_options o;
    int rc = begin_search(g, opt, &o);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g->mu);
    if ((rc = check_batch_sets("mg_sssp_batch_sets", g, row_of, srcs, npairs, k))) return rc;
    if ((rc = sssp_batch_checks("mg_sssp_batch_sets", g, o))) return rc;
    std::vector<int32_t> tsrc((size_t)npairs);
    mg_graph *sg = batch_graph(g, o, srcs, npairs, tsrc.data());
    if (sg != g && (rc = ensure_workspace(sg))) return rc;
    return sssp_batch_run(sg, row_of, sg != g ? tsrc.data() : srcs, npairs, k, dist_out, o, stats, timer);
}
