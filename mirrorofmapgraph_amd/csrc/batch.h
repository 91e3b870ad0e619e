// Shared pieces of the batched searches: 256-bit per-vertex source masks, the
// bitmap walk used by every batched kernel, and the per-graph batch state.
#pragma once
#include "frontier.h"

namespace mg {

constexpr int SB_TASK = 32;  // batched SSSP: edges per wave task
constexpr int SB_RING = 16;  // batched SSSP: bucket bitmaps in the ring
constexpr int SB_PAIR_MAX = 16;  // batched SSSP: vertices with at most this many live slots relax per pair


struct M8 {
    uint4 a, b;
};

__device__ __forceinline__ M8 ld8(const uint4 *p, int v) { return M8{p[2 * v], p[2 * v + 1]}; }
__device__ __forceinline__ void st8(uint4 *p, int v, const M8 &m) {
    p[2 * v] = m.a;
    p[2 * v + 1] = m.b;
}
__device__ __forceinline__ M8 or8(const M8 &x, const M8 &y) {
    return M8{make_uint4(x.a.x | y.a.x, x.a.y | y.a.y, x.a.z | y.a.z, x.a.w | y.a.w),
              make_uint4(x.b.x | y.b.x, x.b.y | y.b.y, x.b.z | y.b.z, x.b.w | y.b.w)};
}
__device__ __forceinline__ M8 andnot8(const M8 &x, const M8 &y) {
    return M8{make_uint4(x.a.x & ~y.a.x, x.a.y & ~y.a.y, x.a.z & ~y.a.z, x.a.w & ~y.a.w),
              make_uint4(x.b.x & ~y.b.x, x.b.y & ~y.b.y, x.b.z & ~y.b.z, x.b.w & ~y.b.w)};
}
__device__ __forceinline__ bool any8(const M8 &x) {
    return (x.a.x | x.a.y | x.a.z | x.a.w | x.b.x | x.b.y | x.b.z | x.b.w) != 0;
}
__device__ __forceinline__ uint32_t word8(const M8 &x, int w) {
    switch (w) {
        case 0: return x.a.x;
        case 1: return x.a.y;
        case 2: return x.a.z;
        case 3: return x.a.w;
        case 4: return x.b.x;
        case 5: return x.b.y;
        case 6: return x.b.z;
        default: return x.b.w;
    }
}
__device__ __forceinline__ int popc8(const M8 &x) {
    return __popc(x.a.x) + __popc(x.a.y) + __popc(x.a.z) + __popc(x.a.w) + __popc(x.b.x) + __popc(x.b.y) +
           __popc(x.b.z) + __popc(x.b.w);
}
__device__ __forceinline__ M8 shfl_xor8(const M8 &x, int o) {
    return M8{make_uint4(__shfl_xor(x.a.x, o), __shfl_xor(x.a.y, o), __shfl_xor(x.a.z, o), __shfl_xor(x.a.w, o)),
              make_uint4(__shfl_xor(x.b.x, o), __shfl_xor(x.b.y, o), __shfl_xor(x.b.z, o), __shfl_xor(x.b.w, o))};
}

// Calls fn(v) for the set bits of a bitmap, block-uniformly: each round a
// workgroup compacts BLK/2 words into LDS and hands them out BLK at a time (fn is
// called by every thread, with v = -1 for padding).
template <class Fn>
__device__ void for_each_bit(const uint32_t *bm, int nwords, Fn fn) {
    constexpr int WPR = BLK / 2;
    __shared__ int s_q[WPR * 32];
    __shared__ int s_w[BLK / WAVE];
    for (int w0 = blockIdx.x * WPR; w0 < nwords; w0 += gridDim.x * WPR) {
        const int w = w0 + (int)threadIdx.x;
        uint32_t word = (threadIdx.x < WPR && w < nwords) ? bm[w] : 0u;
        int total;
        int k = block_exclusive_scan<BLK>(__popc(word), s_w, &total);
        while (word) {
            const int b = __ffs(word) - 1;
            s_q[k++] = w * 32 + b;
            word &= word - 1;
        }
        __syncthreads();
        for (int t0 = 0; t0 < total; t0 += BLK) {
            const int i = t0 + (int)threadIdx.x;
            fn(i < total ? s_q[i] : -1);
        }
        __syncthreads();
    }
}

// As for_each_bit, and every word is cleared once read (each word is read by one
// thread, so nothing else in the launch may read the bitmap).
template <class Fn>
__device__ void for_each_bit_clear(uint32_t *bm, int nwords, Fn fn) {
    constexpr int WPR = BLK / 2;
    __shared__ int s_q[WPR * 32];
    __shared__ int s_w[BLK / WAVE];
    for (int w0 = blockIdx.x * WPR; w0 < nwords; w0 += gridDim.x * WPR) {
        const int w = w0 + (int)threadIdx.x;
        uint32_t word = (threadIdx.x < WPR && w < nwords) ? bm[w] : 0u;
        if (word) bm[w] = 0;
        int total;
        int k = block_exclusive_scan<BLK>(__popc(word), s_w, &total);
        while (word) {
            const int b = __ffs(word) - 1;
            s_q[k++] = w * 32 + b;
            word &= word - 1;
        }
        __syncthreads();
        for (int t0 = 0; t0 < total; t0 += BLK) {
            const int i = t0 + (int)threadIdx.x;
            fn(i < total ? s_q[i] : -1);
        }
        __syncthreads();
    }
}

// Per-slot state of a batched search (device memory).  A slot holds one query
// (source); slots are seeded and retired while the search keeps running.
struct SlotCtl {
    uint32_t assigned[8];                     // slots holding a query
    uint32_t live[8];                         // slots with frontier pairs after the last level
    int start[MG_BATCH_MAX];                  // BFS: level at which the slot was seeded
    long long off[MG_BATCH_MAX];              // SSSP: bucket base when the slot was seeded
    unsigned long long far_hi[MG_BATCH_MAX];  // SSSP: highest bucket the slot has pairs waiting in
    long long te[MG_BATCH_MAX];               // row_stats: TEPS count of the slot's row
    long long nr[MG_BATCH_MAX];               // row_stats: vertices reached
};

// Device state of one batched search context: the masks, the slot rows and the
// level bookkeeping.  mg_bfs_batch / mg_sssp_batch keep one per graph; every
// serving engine owns one.
struct BatchDev {
    bool sssp = false;
    int S = 0;                 // slots (rows)
    int KS = 0;                // SSSP: distance row stride (S rounded up to 16)
    int G = 0;                 // workgroups of the level kernels
    int64_t bytes = 0;         // device bytes held
    Ctl *ctl = nullptr, *h_ctl = nullptr;        // control block, pinned host copy
    SlotCtl *sc = nullptr, *h_sc = nullptr;      // slot state, pinned host copy
    LevelPart *part = nullptr;                   // 3 * MAX_GRID level partials
    uint4 *slot_or = nullptr;                    // MAX_GRID per-workgroup OR of new frontier masks
    uint32_t *act[2] = {nullptr, nullptr};       // active-vertex bitmaps
    uint32_t *ring = nullptr;                    // SSSP bucket ring (SB_RING bitmaps)
    int *hard = nullptr;                         // BFS pull: long in-lists
    ChunkRec *tasks = nullptr;
    int64_t task_cap = 0;      // wave-task records allocated
    int64_t task_need = 0;     // records one level can produce at most (the default capacity)
    int64_t task_floor = 0;    // capacity reached by regrowing (kept for later calls)
    uint4 *V = nullptr, *F = nullptr, *N = nullptr, *FAR = nullptr;
    int *D = nullptr;                            // SSSP distances, n x KS
    int *rows = nullptr;                         // S x n results
    int *rows_out = nullptr;                     // on a degree-ordered twin: rows in the caller's ids
    int2 *pairs = nullptr;                       // SSSP: (vertex, slot) pairs of sparse vertices
    int64_t pair_cap = 0;
    int nwords = 0;
};

// (slot, source) pairs to seed, passed by value as a kernel argument.
struct SeedList {
    int n;
    int slot[MG_BATCH_MAX];
    int src[MG_BATCH_MAX];
};

// Slot ids, passed by value as a kernel argument.
struct SlotList {
    int n;
    int slot[MG_BATCH_MAX];
};

int batch_dev_alloc(BatchDev &b, mg_graph *g, bool sssp, int slots);
void batch_dev_free(BatchDev &b);
// Device bytes batch_dev_alloc would take.
int64_t batch_dev_bytes(const mg_graph *g, bool sssp, int slots);

// Host launchers shared by the batched entry points and the serving engines; all
// enqueue on `st`.
// Clears the context (no slot assigned, all rows of the first `rows_to_clear`
// slots unreached) and sets the level limit and direction parameters.
int bfs_batch_reset(BatchDev &b, mg_graph *g, hipStream_t st, int iter_limit, int alpha, int beta,
                    int rows_to_clear);
// Seeds the listed slots: their old visited bits and rows are cleared first.
int bfs_batch_seed(BatchDev &b, mg_graph *g, hipStream_t st, const SeedList &sl);
// Seeds npairs (slot, source) pairs, several sources per slot allowed (the slot's
// frontier 0 is the set of its sources); the listed slots are cleared first.
int bfs_batch_seed_pairs(BatchDev &b, mg_graph *g, hipStream_t st, const int32_t *slot_of, const int32_t *srcs,
                         int npairs);
// Argument checks of the *_batch_sets entry points: rows in [0, k), sources in
// [0, n), every row with at least one source.
int check_batch_sets(const char *fn, const mg_graph *g, const int32_t *slot_of, const int32_t *srcs, int32_t npairs,
                     int32_t k);
int bfs_batch_step(BatchDev &b, mg_graph *g, hipStream_t st);
int sssp_batch_reset(BatchDev &b, mg_graph *g, hipStream_t st, int iter_limit, long long delta);
// fresh: the context was just reset (distance columns already clear).
// defer: while the search runs, new slots start at the next bucket (serving; see sb_seed)
int sssp_batch_seed(BatchDev &b, mg_graph *g, hipStream_t st, const SeedList &sl, long long delta, bool fresh,
                    bool defer);
int sssp_batch_step(BatchDev &b, mg_graph *g, hipStream_t st, long long delta);
// SSSP: the distance columns of the listed slots into their rows.
int sssp_batch_extract(BatchDev &b, mg_graph *g, hipStream_t st, const SlotList &sl);
// Per-slot OR of the current frontier masks into sc->live.
int batch_live(BatchDev &b, hipStream_t st);
// TEPS count and reached vertices of the listed rows into sc->te / sc->nr (a vertex
// is reached when its entry differs from `unreached`).
int batch_row_stats(BatchDev &b, mg_graph *g, hipStream_t st, const SlotList &sl, int unreached);
// Clears the assigned bits of `mask`.
int batch_release_slots(BatchDev &b, hipStream_t st, const uint32_t mask[8]);

// ---- batched personalized PageRank (ppr.hip)
constexpr int PPR_SLOTS = 64;  // slots per context: one per lane

struct PprDev {
    int n = 0, ntasks = 0, nbig = 0, npart = 0;
    int cur = 0;                          // X[cur] holds the current ranks
    int width = 8;                        // row width W: 8, 16, 32 or 64 columns
    uint64_t live = 0;                    // slots in use
    int64_t bytes = 0;
    float *X[2] = {nullptr, nullptr};     // ranks, n x W (room for n x PPR_SLOTS)
    float *C = nullptr;                   // contributions
    float *part = nullptr;                // partial sums of long in-lists
    float *rows = nullptr;                // PPR_SLOTS x n results
    ChunkRec *tasks = nullptr;
    int *big = nullptr, *src = nullptr;   // long in-lists; per-slot sources (-1 = idle)
};

// out[s * n + v] = rows[s * n + perm[v]] for the first k rows: a twin's batched rows in
// the caller's numbering.
int launch_unpermute_rows(const int *perm, const int *rows, int n, int k, int *out, hipStream_t st);
// Allocates and clears p for graph g; setup copies are ordered on st.
int ppr_dev_alloc(PprDev &p, mg_graph *g, hipStream_t st);
void ppr_dev_free(PprDev &p);
int64_t ppr_dev_bytes(const mg_graph *g);
// Seeds the listed slots (rank column = e_src) after retiring the `retire` slots.
int ppr_seed(PprDev &p, hipStream_t st, const SeedList &sl, const SlotList &retire);
int ppr_iterate(PprDev &p, mg_graph *g, hipStream_t st, float teleport, int iters);
int ppr_extract_rows(PprDev &p, hipStream_t st, const SlotList &sl);
int ppr_iters_of(const mg_options &o);

int check_batch_sources(const char *fn, const mg_graph *g, const int32_t *srcs, int32_t k);
// The graph's own batch context (allocated once at MG_BATCH_MAX slots).
int graph_batch_dev(mg_graph *g, bool sssp, BatchDev **out);
// Sizes the wave-task queue of the graph's batch context for options o (as
// size_chunk_queue: need * max_queue_sizing / 1.5, or the regrown size), and grows it
// 4x after an overflow; both drop the graph's captured step graphs when they resize.
int size_task_queue(mg_graph *g, BatchDev &b, const mg_options &o);
int grow_task_queue(mg_graph *g, BatchDev &b);
// Ends a batched call: device time, overflow check, stats, result delivery.
int finish_batch(mg_graph *g, BatchDev &b, mg_stats *stats, void *out, size_t bytes, bool out_on_device,
                 const CallTimer &timer);

}  // namespace mg
