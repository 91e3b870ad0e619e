// Engine-internal declarations: the device graph store (mg_graph), the pooled
// per-search workspace and the control block the kernels share.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

#include "mg_internal.h"

#define MG_HIP(call)                                                                              \
    do {                                                                                          \
        hipError_t _e = (call);                                                                   \
        if (_e != hipSuccess)                                                                     \
            return mg::set_error(MG_ERR_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                                 __FILE__, __LINE__);                                             \
    } while (0)

namespace mg {

constexpr int BLK = 256;          // threads per workgroup of the multi-workgroup kernels
constexpr int SMALL_BLK = 1024;   // threads of the single-workgroup small-frontier kernel
constexpr int CHUNK_T = 256;      // frontier vertices above this degree are split ...
constexpr int CHUNK = 256;        // ... into edge chunks of this size (one workgroup pass each)
constexpr int SMALL_Q = 4096;     // small-frontier kernel: max frontier vertices
constexpr int SMALL_E = 16384;    // small-frontier kernel: max frontier edges (BFS)
constexpr int BU_INLINE = 32;     // bottom-up: edges a lane scans before handing the vertex to a wave
constexpr int LANE_T = 32;        // lane/wave scheme: a lane walks vertices of up to LANE_T edges ...
constexpr int TASK = 256;         // ... longer lists become TASK-edge tasks, one wave each
constexpr int SMALL_PER = SMALL_Q / SMALL_BLK;        // small-frontier loops: queue entries per thread
constexpr int SMALL_BIG = SMALL_E / (LANE_T + 1) + 1;  // lists > LANE_T in a frontier of <= SMALL_E edges

enum { DIR_TD = 0, DIR_BU = 1 };

constexpr int MAX_GRID = 2048;    // workgroups of a multi-workgroup launch (256 CUs x 8)
constexpr int BIG_GRID = 8192;    // SSSP level and CC kernels: more workgroups hide their gather latency

struct ChunkRec {
    int v, beg, end, pad;
};

// Per-workgroup partial counters of one level (summed by the level close).
struct LevelPart {
    long long nf;   // vertices entering the next frontier
    long long mf;   // their out-edges
    long long ex;   // edges examined
    long long pad;
};

// Control block shared by the kernels of one search (device memory).  Levels are
// advanced on the device; the host only reads `done` every few queued levels.
struct Ctl {
    int level;          // level being processed
    int done;           // 1 once the search has converged or hit iter_limit
    int dir;            // DIR_TD or DIR_BU for the current level
    int steps;          // BFS / SSSP: steps whose first kernel found the search running
    int qsize[3];       // sizes of the three rotating frontier queues
    int nchunks;        // edge chunks queued this level
    int nhard;          // bottom-up vertices handed to whole waves this level
    int iter_limit;     // stop after this many levels
    int n;
    int can_bu;         // bottom-up allowed (an in-edge view exists)
    int alpha, beta;    // direction heuristic parameters
    int nf_prev;        // frontier size of the previous level
    int bu_levels;
    int small_levels;
    int changed;        // generic "something changed" counter / overflow flag (bit 30)
    int pending;        // a multi-workgroup level finished and awaits its close
    int qvalid;         // bit i: queue q[i] lists the frontier of bitmap bm[i]
    int aux;            // search-specific (CC: root of the sampled large component)
    long long mf_next;        // sum of out-degrees of vertices found this level
    long long m_unexplored;   // out-edges of unreached vertices
    long long edges_traversed;// reference TEPS count
    long long reached;
    long long edges_examined;
    double dsum;              // scratch accumulator (PageRank L1 change)
    long long mf_cur;         // out-edges of the current frontier
    long long thr;            // bucketed SSSP: distances below thr are in the current bucket
    long long far_hi;         // bucketed SSSP: highest bucket holding a waiting distance
    int pair_end;             // batched SSSP: pair slots below this are valid this iteration
    long long fmin;           // SSSP: a lower bound on the distances of the current frontier
};

// Pooled per-search device state, grown on demand and reused across searches.
struct Workspace {
    int n = 0;
    int nwords = 0;           // 32-bit words of one vertex bitmap (rounded to even)
    int64_t chunk_cap = 0;    // edge-chunk records allocated
    int64_t chunk_need = 0;   // records one level can produce at most (the default capacity)
    int64_t chunk_floor = 0;  // capacity reached by regrowing (kept for later searches)
    int32_t *vals = nullptr;  // depth / dist / label
    long long *trace = nullptr;  // per level: next frontier size | direction << 40, its out-edges
    int32_t *q[3] = {nullptr, nullptr, nullptr};
    int32_t *hard = nullptr;
    uint32_t *visited = nullptr;
    uint32_t *bm[3] = {nullptr, nullptr, nullptr};
    ChunkRec *chunks = nullptr;
    LevelPart *part = nullptr;  // 3 * BIG_GRID slots
    Ctl *ctl = nullptr;
    Ctl *h_ctl = nullptr;     // pinned host mirror
    unsigned long long *h_reach = nullptr;  // pinned: TEPS count and reached vertices of a result
    unsigned long long *d_reach = nullptr;  // their device accumulators
    // PageRank state (allocated on first use)
    float *rank = nullptr;
    float *contrib[2] = {nullptr, nullptr};
    double *partials = nullptr;                       // per-workgroup L1 partial sums
    int32_t *pr_mid = nullptr;                        // vertices of medium in-degree (a wave each)
    ChunkRec *pr_chunks = nullptr;                    // in-edge chunks of high in-degree vertices
    double *pr_partials = nullptr;                    // per-workgroup partials of the PageRank gather
    LevelPart *pr_cpart = nullptr;
    int pr_nslots = 0;
    int32_t *pr_bigv = nullptr;                       // first chunk of each high in-degree vertex (+ end)
    float *pr_part = nullptr;                         // per-chunk partial sums
    int32_t *pr_pact = nullptr;                       // per-chunk "an in-neighbour changed" flags
    int pr_nmid = 0, pr_nchunks = 0, pr_nbig = 0;
    const int32_t *pr_view = nullptr;                 // in-edge offsets the buckets were built for
    // Static pull tasks: PULL_TASK-edge pieces of the in-edge lists longer than LANE_T
    ChunkRec *pull_tasks = nullptr;
    int pull_ntasks = 0;
    const int32_t *pull_view = nullptr;
    ChunkRec *dense_tasks = nullptr;  // dense SSSP pull tasks (pieces and vertex ranges)
    int dense_ntasks = 0;
    const int32_t *dense_view = nullptr;
};

}  // namespace mg

struct mg_graph {
    int device = 0;
    hipStream_t stream = nullptr;
    bool stream_borrowed = false;  // a twin runs on its parent's stream
    bool arrays_borrowed = false;  // a search view: the device arrays belong to another graph
    bool no_capture = false;       // steps run as direct launches (no captured step graphs)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int32_t n = 0;
    int64_t m = 0;
    uint32_t flags = 0;
    int32_t *row_off = nullptr, *col = nullptr, *w = nullptr;
    int32_t *col_off = nullptr, *row_idx = nullptr, *csc_eid = nullptr;
    int32_t min_weight = 0;
    int32_t max_weight = 0;
    double mean_weight = 0;
    int64_t topo_bytes = 0;
    std::mutex mu;
    mg::Workspace ws;
    struct mg_graph_steps *steps = nullptr;  // captured step graphs
    // MG_GRAPH_RELABEL: `twin` is the graph renamed in out-degree order (owned; used under
    // this graph's mutex) and h_perm[v] its id for vertex v.  On the twin, perm is the
    // device copy of that map and unperm receives a result in the caller's numbering.
    mg_graph *twin = nullptr;
    std::vector<int32_t> h_perm;
    int32_t *perm = nullptr;
    int32_t *unperm = nullptr;
};

namespace mg {

// A copy ordered on stream st that returns when it is done.  Library code never uses
// the blocking null-stream hipMemcpy: that waits on every blocking stream of the device
// and fails while another host thread captures a step graph.
inline hipError_t copy_on(hipStream_t st, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    if (!bytes) return hipSuccess;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
    return e == hipSuccess ? hipStreamSynchronize(st) : e;
}

// The graph a BFS / SSSP / PageRank search with options o runs on (the degree-ordered
// twin when there is one and o.layout allows it) and the source's id there.  Results
// of a search on a twin are mapped back to the caller's numbering by finish_search.
inline mg_graph *search_graph(mg_graph *g, const mg_options &o, int *src) {
    if (!g->twin || o.layout != 0 || o.engine != MG_ENGINE_AUTO) return g;
    if (src) *src = g->h_perm[(size_t)*src];
    return g->twin;
}

// As search_graph for a batched call: the graph it runs on, with the sources renamed
// into out[0, k) when that is the twin.
inline mg_graph *batch_graph(mg_graph *g, const mg_options &o, const int32_t *srcs, int k, int32_t *out) {
    mg_graph *t = search_graph(g, o, nullptr);
    if (t != g)
        for (int i = 0; i < k; ++i) out[i] = g->h_perm[(size_t)srcs[i]];
    return t;
}

// Grows the workspace for graph g; returns an MG_* status.
int ensure_workspace(mg_graph *g);
// Fault injection for tests: with MG_FAIL_ALLOC_AT=k in the environment, the k-th
// device allocation of the library counted from the moment the variable was set (or
// last changed) is refused, so every cleanup path behind an MG_ERR_OOM can be driven
// from the host.  Each allocation site calls this first; false when unset.
bool alloc_refused();
// Queue sizing.  A queue whose records one level can produce at most `need` holds
// need * max_queue_sizing / 1.5 records (the reference's max_queue_sizing, default
// 1.5, scales its frontier queues the same way, csr_problem.cuh:627-827), at least
// 64; a queue that overflows is regrown 4x and the search replayed, and the grown
// size stays for later searches on the graph.
constexpr int MAX_REGROWS = 12;
int64_t queue_capacity(int64_t need, double max_queue_sizing);
// Sizes the edge-chunk queue of the BFS / SSSP levels for options o (reallocating it,
// and dropping the captured step graphs that point at it, when the size changes).
int size_chunk_queue(mg_graph *g, const mg_options &o);
// After an overflow: 4x the capacity (kept for later searches).
int grow_chunk_queue(mg_graph *g);
// Destroys every captured step graph of g (their kernel arguments hold buffer
// pointers; they are captured again on next use).
void invalidate_steps(mg_graph *g);
int ensure_pr_workspace(mg_graph *g);
int alloc_i32(int32_t **p, size_t n);
int alloc_f32(float **p, size_t n);
int alloc_chunks(ChunkRec **p, size_t n);
int alloc_f64(double **p, size_t n);
int alloc_parts(LevelPart **p, size_t n);
void free_dev(void *p);
int grid_blocks(int64_t work_items, int block);
// As grid_blocks, capped at `cap` workgroups instead of MAX_GRID.
int grid_blocks_cap(int64_t work_items, int block, int cap);

// Per-call bookkeeping shared by the search entry points.
struct CallTimer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

// Validates the common search arguments and locks the graph.
int begin_search(mg_graph *g, const mg_options *opt, mg_options *resolved);
// Copies a device int32/float result to the caller (host or device pointer).
int deliver(mg_graph *g, void *dst, const void *src, size_t bytes, bool out_on_device);
// Polls ctl->done on the host every `k` queued steps; `step` enqueues one step.
// A captured run of STEPS_PER_GRAPH consecutive steps of one search kind.  Kernel
// arguments are fixed per (graph, kind, key), so the executable graph is reused by
// every later search with the same key.
constexpr int STEPS_PER_GRAPH = 8;
constexpr int TRACE_CAP = 16384;  // levels recorded in the per-level trace (verbose)
constexpr int PULL_TASK = 512;   // static pull tasks: in-edges per wave (8 per lane in flight)
// Dense SSSP pull tasks: in-lists longer than DENSE_BIG are cut into DENSE_T-slot
// pieces; the other vertices are grouped into ranges of whole bitmap words (at most
// DENSE_V vertices) holding about DENSE_T in-edges.  A range holds at most DENSE_T
// in-edges, or a single word's (32 vertices of up to DENSE_BIG each): DENSE_RANGE_MAX.
constexpr int DENSE_BIG = 512;
constexpr int DENSE_T = 2048;
constexpr int DENSE_V = 256;
constexpr int DENSE_RANGE_MAX = DENSE_T > 32 * DENSE_BIG ? DENSE_T : 32 * DENSE_BIG;
enum { KIND_BFS = 0, KIND_SSSP = 1, KIND_PR = 2, KIND_GAS = 3, KIND_BATCH = 4, KIND_BATCH_SSSP = 5, NKINDS = 8 };

struct StepGraph {
    hipGraphExec_t exec = nullptr;                        // STEPS_PER_GRAPH steps
    hipGraphExec_t part[STEPS_PER_GRAPH] = {};  // part[i]: i steps (1 <= i < STEPS_PER_GRAPH; captured on first use)
    uint64_t key = ~0ULL;
    int64_t last_steps = 0;  // steps the previous search with this key ran (0: not counted)
};

// Returns MG_OK once the device sets ctl->done; a search still running after
// max_steps steps is an error (every iteration limit sets done on the device, so
// this never ends a normal search early).
struct NoTail {
    int operator()() const { return 0; }
};
// tail (optional) queues the work that follows a finished search.  When the step count
// is known from the previous search it is queued behind those steps, before the one
// wait; *tail_queued then says whether the search finished there with the tail run.
template <class Step, class Tail = NoTail>
int drive(mg_graph *g, Step step, int64_t max_steps, int kind = -1, uint64_t key = 0, Ctl *ctl = nullptr,
          Ctl *h_ctl = nullptr, Tail tail = Tail(), bool *tail_queued = nullptr);
inline int not_converged(int64_t max_steps) {
    return set_error(MG_ERR_DEVICE, "the search did not finish within %lld steps", (long long)max_steps);
}
// Captured step graphs are on unless MG_NO_GRAPH is set to a non-zero value (then
// every step is launched directly; read once per process).
bool graphs_enabled();
// A stream on a hardware queue of its own: streams created with a CU mask (here every
// CU) are not pooled with the process's other streams, so a long kernel queued on
// another stream never sits ahead of this one's work.  Such a stream synchronises with
// the null stream (it is not created non-blocking), so no step graph is captured on it.
hipError_t dedicated_stream(int device, hipStream_t *st);
// A search view of g: g's device arrays (not owned; g must outlive the view) with its
// own dedicated stream, events, lock, workspace and steps (direct launches), and a view
// of g's twin on the same stream.  Searches on the view give g's results; they run
// beside searches on g and on other views.
int graph_view_create(mg_graph *g, mg_graph **out);

// Generic GAS enactor over the ported programs (engine = MG_ENGINE_GAS).
struct GasStateBase;
struct BatchDev;
struct PprDev;
int gas_bfs(mg_graph *g, int src, int *out, const mg_options &o, mg_stats *st, const CallTimer &t);
int gas_sssp(mg_graph *g, int src, int *out, const mg_options &o, mg_stats *st, const CallTimer &t);
int gas_cc(mg_graph *g, int *out, const mg_options &o, mg_stats *st, const CallTimer &t);
int gas_pagerank(mg_graph *g, float *out, const mg_options &o, mg_stats *st, const CallTimer &t);
void gas_release(mg_graph *g);
void batch_release(mg_graph *g);
void ppr_release(mg_graph *g);
int64_t batch_dev_size(const BatchDev *b);  // device bytes held by a batch context
int64_t ppr_dev_size(const PprDev *p);
// SSSP bucket width for options o on graph g (a huge value when unbucketed).
long long sssp_delta(const mg_graph *g, const mg_options &o);

}  // namespace mg

struct mg_graph_steps {
    mg::StepGraph sg[mg::NKINDS];
    mg::GasStateBase *gas[4] = {nullptr, nullptr, nullptr, nullptr};
    mg::BatchDev *bdev[2] = {nullptr, nullptr};  // batched BFS / SSSP contexts
    mg::PprDev *pdev = nullptr;                  // batched personalized PageRank context
};

// Queues steps until the device reports `done`.  Steps run from a captured HIP graph
// (kind >= 0) or as direct launches; the host reads the control block only between
// batches of steps (1, 2, 4, ... graphs), never once per level.
template <class Step, class Tail>
int mg::drive(mg_graph *g, Step step, int64_t max_steps, int kind, uint64_t key, Ctl *ctl, Ctl *h_ctl, Tail tail,
              bool *tail_queued) {
    Workspace &ws = g->ws;
    if (tail_queued) *tail_queued = false;
    if (!ctl) {
        ctl = ws.ctl;
        h_ctl = ws.h_ctl;
    }
    if (kind >= 0 && kind < NKINDS && graphs_enabled() && !g->no_capture) {
        StepGraph &sg = g->steps->sg[kind];
        auto capture = [&](int nsteps, hipGraphExec_t *out) -> int {
            MG_HIP(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
            int rc = MG_OK;
            for (int i = 0; i < nsteps && rc == MG_OK; ++i) rc = step();
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(g->stream, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (e != hipSuccess) return set_error(MG_ERR_DEVICE, "stream capture failed: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                *out = nullptr;
                return set_error(MG_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(e));
            }
            return MG_OK;
        };
        if (!sg.exec || sg.key != key) {
            if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
            for (hipGraphExec_t &x : sg.part) {
                if (x) (void)hipGraphExecDestroy(x);
                x = nullptr;
            }
            sg.exec = nullptr;
            sg.last_steps = 0;
            int rc = capture(STEPS_PER_GRAPH, &sg.exec);
            if (rc) return rc;
            sg.key = key;
        }
        auto finish = [&]() {
            // kinds whose steps count themselves: the next search starts with this many
            sg.last_steps = (kind == KIND_BFS || kind == KIND_SSSP) ? h_ctl->steps : 0;
        };
        int64_t queued = 0, k = 1;
        // A search like the previous one (same graph, kind and options) ran last_steps
        // steps: queue exactly that many (full graphs, then one graph of the rest) before
        // the first look at the control block, so no empty steps trail the search.
        if (sg.last_steps > 0) {
            int64_t left = std::min<int64_t>(sg.last_steps, max_steps);
            for (; left >= STEPS_PER_GRAPH; left -= STEPS_PER_GRAPH, queued += STEPS_PER_GRAPH)
                MG_HIP(hipGraphLaunch(sg.exec, g->stream));
            if (left > 0) {  // one graph of the remaining steps (each graph boundary costs ~10 us)
                if (!sg.part[left]) {
                    int rc = capture((int)left, &sg.part[left]);
                    if (rc) return rc;
                }
                MG_HIP(hipGraphLaunch(sg.part[left], g->stream));
                queued += left;
            }
            // the search most likely ends here: queue its tail before the wait (if not,
            // the caller queues it again after the remaining steps)
            if (tail_queued) {
                const int trc = tail();
                if (trc) return trc;
            }
            MG_HIP(hipMemcpyAsync(h_ctl, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, g->stream));
            MG_HIP(hipStreamSynchronize(g->stream));
            if (h_ctl->done) {
                finish();
                if (tail_queued) *tail_queued = true;
                return MG_OK;
            }
        }
        while (queued < max_steps) {
            for (int i = 0; i < k && queued < max_steps; ++i, queued += STEPS_PER_GRAPH)
                MG_HIP(hipGraphLaunch(sg.exec, g->stream));
            MG_HIP(hipMemcpyAsync(h_ctl, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, g->stream));
            MG_HIP(hipStreamSynchronize(g->stream));
            if (h_ctl->done) {
                finish();
                return MG_OK;
            }
            if (k < 16) k *= 2;
        }
        return not_converged(max_steps);
    }
    int64_t queued = 0, k = 2;
    while (queued < max_steps) {
        for (int i = 0; i < k && queued < max_steps; ++i, ++queued) {
            int rc = step();
            if (rc) return rc;
        }
        MG_HIP(hipMemcpyAsync(h_ctl, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, g->stream));
        MG_HIP(hipStreamSynchronize(g->stream));
        if (h_ctl->done) return MG_OK;
        if (k < 64) k *= 2;
    }
    return not_converged(max_steps);
}
