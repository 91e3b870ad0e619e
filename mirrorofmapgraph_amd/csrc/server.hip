// Serving: traversal queries on resident graphs, with continuous batching.
//
// Not in the reference (each of its drivers runs one search at a time, e.g.
// Algorithms/BFS/bfs.cu:402-417).  Queries (BFS, SSSP, personalized PageRank) on
// graphs added to the server are served by one engine per (graph, kind).  An
// engine owns a batch context with up to 256 slots (64 for PageRank) and its own
// HIP stream, so engines of different graphs and kinds overlap on the device.
// Engines advance in rounds driven by one dispatcher thread; a round
//   1. retires finished slots: their rows are copied out (SSSP: transposed out of
//      the vertex-major distances first) with per-slot TEPS counts;
//   2. admits waiting queries into free slots: seeded into the running search, which
//      does not stop (a BFS slot joins the next level, an SSSP slot the current
//      bucket with its own offset, a PageRank slot the next iteration);
//   3. runs a chunk of steps (a captured HIP graph of STEPS_PER_GRAPH levels; a few
//      PageRank iterations) and copies the slot status back;
// and ends with an event the dispatcher polls, so no round waits for another
// engine.  Status blocks and result rows leave through copy kernels on the engine's
// own stream, not DMA copies: a DMA copy queued behind a long kernel waits inside a
// DMA queue that every stream's copies share, so one thin search in flight (tens of
// milliseconds) held up every engine's status copy behind it.  An engine's device state is allocated when it first has work, within
// the server's HBM budget (idle engines give theirs back when another needs it;
// an engine that cannot fit even one slot leaves its queries waiting).  Results do
// not depend on what else runs: BFS depths and SSSP distances are unique, and a
// PageRank slot's arithmetic is fixed by the graph.
#include "batch.h"
#include "thin.h"

#include <atomic>
#include <climits>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

namespace mg {

struct Query {
    int64_t id = 0;
    int graph = 0, kind = 0, src = 0;
    void *out = nullptr;
    void *out_dev = nullptr;  // out as the device addresses it (null: host memory the device cannot write)
    bool out_on_device = false;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int status = MG_OK;
    bool done = false;
    int slot = -1;
    int iters_left = 0;  // PageRank
    int64_t te = 0, nr = 0;
};

struct Server;

// A launch of thin searches in flight on its own stream.
struct ThinSlot {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    std::vector<Query *> qs;
};

struct Engine {
    Server *srv = nullptr;
    mg_graph *g = nullptr;                  // the graph the engine searches: the registered
                                            // graph's degree-ordered twin when it has one
                                            // (batched engines), else the graph itself
    const int32_t *h_perm = nullptr;        // caller id -> twin id (host), null without a twin
    const int *d_perm = nullptr;            // the same map on the device
    int kind = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    bool allocated = false, inflight = false;
    int S = 0;                              // slots
    BatchDev b;                             // BFS / SSSP
    Ctl *h_ctl_dev = nullptr;               // b.h_ctl / b.h_sc as the device addresses them
    SlotCtl *h_sc_dev = nullptr;
    PprDev p;                               // PageRank
    hipGraphExec_t chunk = nullptr;         // captured steps
    long long delta = 1;                    // SSSP bucket width
    std::deque<Query *> waiting;
    std::vector<Query *> slot_q;            // slot -> running query
    std::vector<Query *> retiring;          // rows copied in the round in flight
    int running = 0;
    double idle_since = 0;
    // plain engines: one whole-GPU search per query (mg_bfs / mg_sssp) on a worker thread
    bool plain = false;
    std::vector<mg_graph *> views;          // a plain engine's search views (owned), one per worker
    std::vector<std::thread> workers;
    std::mutex wmu;                         // the workers' wake-up (wake_gen / quit under wmu)
    std::condition_variable wcv;
    uint64_t wake_gen = 0;
    bool quit = false;
    // thin engines (BFS / SSSP on a high-diameter graph): one workgroup per query
    bool thin = false;
    ThinDev td;
    std::vector<ThinSlot> tslots;
    std::vector<int> free_lanes;
    std::deque<Query *> paused;             // thin searches waiting to be launched again
    int64_t bytes() const { return allocated ? (kind == MG_Q_PPR ? p.bytes : thin ? td.bytes : b.bytes) : 0; }
};

struct Server {
    mg_server_config cfg{};
    long long thin_slice = 0;  // wall-clock ticks of a thin launch's slice (0: none)
    std::vector<mg_graph *> graphs;
    std::vector<std::unique_ptr<Engine>> engines;  // graph * 3 + kind
    std::unordered_map<int64_t, std::unique_ptr<Query>> queries;
    // mu guards the queues and query states.  The dispatcher holds it through its
    // passes; other threads announce themselves in `want` (fair_lock) and the
    // dispatcher steps aside between passes while any waits.  Waiting for a query
    // does not sleep on mu: finished queries bump done_seq under dmu and wake dcv.
    std::mutex mu;
    std::atomic<int> want{0};
    std::condition_variable cv_work;
    std::mutex dmu;
    std::condition_variable dcv;
    std::atomic<uint64_t> done_seq{0};
    std::thread th;
    bool stop = false;
    int device = 0;
    int64_t next_id = 1;
    int64_t bytes_used = 0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    mg_server_stats stats{};
    int error = MG_OK;
    double now() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

// dst[i] = src[i]: device memory into device memory or pinned host memory.
__global__ void copy_out_v4(const uint4 *src, uint4 *dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
__global__ void copy_out_w(const uint32_t *src, uint32_t *dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Queues a kernel copy of `bytes` (a multiple of 4) on st; dst must be addressable by
// the device (device memory, or pinned host memory through its device address).
static int copy_out(void *dst, const void *src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return MG_OK;
    if (((uintptr_t)dst | (uintptr_t)src | bytes) % 16 == 0) {
        const int64_t n = (int64_t)(bytes / 16);
        hipLaunchKernelGGL(copy_out_v4, dim3(grid_blocks_cap(n, BLK, 512)), dim3(BLK), 0, st, (const uint4 *)src,
                           (uint4 *)dst, n);
    } else {
        const int64_t n = (int64_t)(bytes / 4);
        hipLaunchKernelGGL(copy_out_w, dim3(grid_blocks_cap(n, BLK, 512)), dim3(BLK), 0, st, (const uint32_t *)src,
                           (uint32_t *)dst, n);
    }
    MG_HIP(hipGetLastError());
    return MG_OK;
}

// A finished query's row into its output: a kernel copy where the device can write
// the output, else a runtime copy (host memory the runtime stages; it blocks the
// dispatcher while it copies, as documented for mg_server_submit).
static int deliver_row(const Query *r, const void *row, size_t bytes, hipStream_t st) {
    if (r->out_dev) return copy_out(r->out_dev, row, bytes, st);
    if (r->out) MG_HIP(hipMemcpyAsync(r->out, row, bytes, hipMemcpyDeviceToHost, st));
    return MG_OK;
}

// The same for a row of an engine that searches a twin: out[v] = row[d_perm[v]], by a
// gather kernel straight into the output where the device can write it, else through
// the host (the documented blocking path for other host memory).
static int deliver_twin_row(const Query *r, const int *d_perm, const int32_t *h_perm, const void *row, int n,
                            hipStream_t st) {
    if (r->out_dev) return launch_unpermute(d_perm, row, n, r->out_dev, st);
    if (!r->out) return MG_OK;
    std::vector<int32_t> tmp((size_t)n);
    MG_HIP(hipMemcpyAsync(tmp.data(), row, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
    int32_t *o = (int32_t *)r->out;
    for (int v = 0; v < n; ++v) o[v] = tmp[(size_t)h_perm[v]];
    return MG_OK;
}

// The device address of out: itself for device memory, the mapped address of pinned
// host memory, null for host memory the device cannot write.
static void *device_view(void *out, bool on_device) {
    if (!out || on_device) return out;
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, out) != hipSuccess) {
        (void)hipGetLastError();  // not memory the runtime knows: pageable host memory
        return nullptr;
    }
    return at.type == hipMemoryTypeHost ? at.devicePointer : nullptr;
}

static int engine_steps(Engine &e) {
    return e.srv->cfg.steps_per_round > 0 ? e.srv->cfg.steps_per_round : STEPS_PER_GRAPH;
}

static void engine_free(Server &s, Engine &e) {
    if (!e.allocated) return;
    if (e.chunk) (void)hipGraphExecDestroy(e.chunk);
    e.chunk = nullptr;
    s.bytes_used -= e.bytes();
    if (e.kind == MG_Q_PPR) ppr_dev_free(e.p);
    else if (e.thin) thin_dev_free(e.td);
    else batch_dev_free(e.b);
    e.h_ctl_dev = nullptr;
    e.h_sc_dev = nullptr;
    e.free_lanes.clear();
    e.allocated = false;
    e.S = 0;
}

// Allocates the engine's device state within the budget: idle engines are freed
// first if needed, and the slot count is halved until it fits.
static int thin_lanes(const Server &s) { return s.cfg.thin_lanes > 0 ? s.cfg.thin_lanes : 64; }
// Thin launches in flight per engine.
static int thin_streams(const Server &s) { return s.cfg.thin_streams > 0 ? s.cfg.thin_streams : 1; }

static int engine_alloc(Server &s, Engine &e) {
    if (e.thin) {
        const bool sssp = e.kind == MG_Q_SSSP;
        int L = thin_lanes(s);
        if (s.bytes_used + thin_dev_bytes(e.g, sssp, L) > s.cfg.hbm_budget) {
            for (auto &o : s.engines)
                if (o && o.get() != &e && o->allocated && !o->inflight && o->running == 0 && o->waiting.empty())
                    engine_free(s, *o);
        }
        while (L > 1 && s.bytes_used + thin_dev_bytes(e.g, sssp, L) > s.cfg.hbm_budget) L /= 2;
        if (s.bytes_used + thin_dev_bytes(e.g, sssp, L) > s.cfg.hbm_budget) return MG_ERR_OOM;  // wait for memory
        mg_options o;
        mg_options_default(&o);
        o.sssp_delta = s.cfg.sssp_delta;
        e.delta = sssp_delta(e.g, o);
        int rc = thin_dev_alloc(e.td, e.g, sssp, L, e.delta);
        if (rc) return rc;
        e.allocated = true;
        s.bytes_used += e.bytes();
        e.free_lanes.clear();
        for (int l = L - 1; l >= 0; --l) e.free_lanes.push_back(l);
        return MG_OK;
    }
    const int want = e.kind == MG_Q_PPR ? PPR_SLOTS : std::min(MG_BATCH_MAX, std::max(1, s.cfg.max_slots));
    auto need = [&](int S) {
        return e.kind == MG_Q_PPR ? ppr_dev_bytes(e.g) : batch_dev_bytes(e.g, e.kind == MG_Q_SSSP, S);
    };
    int S = want;
    if (s.bytes_used + need(S) > s.cfg.hbm_budget) {
        for (auto &o : s.engines)
            if (o && o.get() != &e && o->allocated && !o->inflight && o->running == 0 && o->waiting.empty())
                engine_free(s, *o);
    }
    while (S > 1 && s.bytes_used + need(S) > s.cfg.hbm_budget) S = e.kind == MG_Q_PPR ? 0 : S / 2;
    if (S < 1 || s.bytes_used + need(S) > s.cfg.hbm_budget) return MG_ERR_OOM;  // wait for memory
    int rc;
    if (e.kind == MG_Q_PPR) {
        rc = ppr_dev_alloc(e.p, e.g, e.st);
    } else {
        rc = batch_dev_alloc(e.b, e.g, e.kind == MG_Q_SSSP, S);
        if (rc == MG_OK && (hipHostGetDevicePointer((void **)&e.h_ctl_dev, e.b.h_ctl, 0) != hipSuccess ||
                            hipHostGetDevicePointer((void **)&e.h_sc_dev, e.b.h_sc, 0) != hipSuccess)) {
            batch_dev_free(e.b);
            return set_error(MG_ERR_DEVICE, "pinned status blocks are not mapped for the device");
        }
    }
    if (rc) return rc;
    e.S = S;
    e.allocated = true;
    s.bytes_used += e.bytes();
    e.slot_q.assign(S, nullptr);
    if (e.kind == MG_Q_BFS) {
        if ((rc = bfs_batch_reset(e.b, e.g, e.st, INT_MAX, 4, 16, 0))) return rc;
    } else if (e.kind == MG_Q_SSSP) {
        mg_options o;
        mg_options_default(&o);
        o.sssp_delta = s.cfg.sssp_delta;
        e.delta = sssp_delta(e.g, o);
        if ((rc = sssp_batch_reset(e.b, e.g, e.st, INT_MAX, e.delta))) return rc;
    }
    if (e.kind != MG_Q_PPR && graphs_enabled()) {
        MG_HIP(hipStreamBeginCapture(e.st, hipStreamCaptureModeThreadLocal));
        int src = MG_OK;
        for (int i = 0; i < engine_steps(e) && src == MG_OK; ++i)
            src = e.kind == MG_Q_BFS ? bfs_batch_step(e.b, e.g, e.st) : sssp_batch_step(e.b, e.g, e.st, e.delta);
        if (src == MG_OK) src = batch_live(e.b, e.st);
        hipGraph_t graph = nullptr;
        hipError_t he = hipStreamEndCapture(e.st, &graph);
        if (src) {
            if (graph) (void)hipGraphDestroy(graph);
            return src;
        }
        if (he != hipSuccess) return set_error(MG_ERR_DEVICE, "step capture failed: %s", hipGetErrorString(he));
        he = hipGraphInstantiate(&e.chunk, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (he != hipSuccess) return set_error(MG_ERR_DEVICE, "step graph instantiation failed: %s", hipGetErrorString(he));
    }
    return MG_OK;
}

static void finish_query(Server &s, Query *q, double t) {
    q->t_done = t;
    q->done = true;
    s.stats.completed++;
    s.stats.edges_traversed += q->te;
    s.stats.latency_ms_sum += q->t_done - q->t_submit;
    {
        std::lock_guard<std::mutex> g(s.dmu);
        s.done_seq.fetch_add(1);
    }
    s.dcv.notify_all();
}

// mu for a thread other than the dispatcher (see Server::want).
static std::unique_lock<std::mutex> fair_lock(Server &s) {
    s.want.fetch_add(1);
    std::unique_lock<std::mutex> lk(s.mu);
    s.want.fetch_sub(1);
    return lk;
}

// One round of a thin engine: finished launches hand their queries back (rows and
// counts are in place when a launch's event completes), then waiting queries are
// launched, one workgroup each, on every idle launch stream.  Called with s.mu held.
static int thin_round(Server &s, Engine &e, bool *queued) {
    *queued = false;
    for (ThinSlot &ts : e.tslots) {
        if (!ts.busy) continue;
        const hipError_t q = hipEventQuery(ts.ev);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) return set_error(MG_ERR_DEVICE, "thin search event: %s", hipGetErrorString(q));
        ts.busy = false;
        const double t = s.now();
        for (Query *r : ts.qs) {
            const long long *hs = e.td.h_stats + THIN_STAT * r->slot;
            if (hs[2] == 0) {  // paused at the end of its slice: launched again next
                e.paused.push_back(r);
                continue;
            }
            if (r->out && !r->out_dev) {  // host memory the device cannot write: a runtime copy
                MG_HIP(hipMemcpyAsync(r->out, e.td.a.rows + (size_t)r->slot * e.g->n, 4 * (size_t)e.g->n,
                                      hipMemcpyDeviceToHost, ts.st));
                MG_HIP(hipStreamSynchronize(ts.st));
            }
            r->te = hs[0];
            r->nr = hs[1];
            e.free_lanes.push_back(r->slot);
            r->slot = -1;
            finish_query(s, r, t);
            e.running--;
        }
        ts.qs.clear();
    }
    e.inflight = false;
    for (const ThinSlot &ts : e.tslots) e.inflight |= ts.busy;
    if (e.waiting.empty() && e.paused.empty()) {
        if (e.running == 0 && e.idle_since == 0) e.idle_since = s.now();
        return MG_OK;
    }
    e.idle_since = 0;
    if (!e.allocated) {
        const int rc = engine_alloc(s, e);
        if (rc == MG_ERR_OOM && !e.allocated) {
            clear_error();
            return MG_OK;  // queries wait for memory
        }
        if (rc) return rc;
    }
    for (ThinSlot &ts : e.tslots) {
        if (ts.busy || (e.paused.empty() && (e.waiting.empty() || e.free_lanes.empty()))) continue;
        ThinBatch jb{};
        jb.slice = s.thin_slice;
        // paused searches first (oldest first), then new queries on free lanes
        while (jb.n < THIN_MAX_JOBS && !e.paused.empty()) {
            Query *r = e.paused.front();
            e.paused.pop_front();
            jb.lane[jb.n] = r->slot;
            jb.src[jb.n] = -1;
            jb.out[jb.n++] = r->out_dev;
            ts.qs.push_back(r);
        }
        while (jb.n < THIN_MAX_JOBS && !e.waiting.empty() && !e.free_lanes.empty()) {
            Query *r = e.waiting.front();
            e.waiting.pop_front();
            r->slot = e.free_lanes.back();
            e.free_lanes.pop_back();
            r->t_admit = s.now();
            jb.lane[jb.n] = r->slot;
            jb.src[jb.n] = r->src;
            jb.out[jb.n++] = r->out_dev;
            ts.qs.push_back(r);
            e.running++;
        }
        const int rc = thin_launch(e.td.a, jb, e.kind == MG_Q_SSSP, ts.st);
        if (rc) return rc;
        MG_HIP(hipEventRecord(ts.ev, ts.st));
        ts.busy = true;
        e.inflight = true;
        *queued = true;
        s.stats.rounds++;
    }
    return MG_OK;
}

// PageRank cohorts.  A sweep over the graph costs the same for one live slot as for
// eight (rows are at least 8 wide), so serving each query on arrival keeps a lightly
// loaded engine sweeping nonstop for one or two queries at a time.  An idle engine
// starts when ppr_batch_min queries wait or the oldest has waited ppr_gate_us, and a
// running cohort takes new queries during its first PPR_JOIN iterations only; the
// others start the next cohort.
constexpr int PPR_JOIN = 8;
static bool ppr_admits(const Server &s, const Engine &e) {
    if (s.cfg.ppr_gate_us < 0 || e.waiting.empty()) return true;
    if (e.running > 0) {
        int most = 0;
        for (const Query *r : e.slot_q)
            if (r) most = std::max(most, s.cfg.ppr_iters - r->iters_left);
        return most < PPR_JOIN;
    }
    const int batch = s.cfg.ppr_batch_min > 0 ? s.cfg.ppr_batch_min : 8;
    const double gate_ms = (s.cfg.ppr_gate_us > 0 ? s.cfg.ppr_gate_us : 20000) / 1000.0;
    return (int)e.waiting.size() >= batch || s.now() - e.waiting.front()->t_submit >= gate_ms;
}

// One round of an engine; returns true if it queued work.  Called with s.mu held.
static int engine_round(Server &s, Engine &e, bool *queued) {
    *queued = false;
    if (e.plain) return MG_OK;  // served by its worker thread
    if (e.thin) return thin_round(s, e, queued);
    *queued = false;
    std::vector<int> finished;
    if (e.inflight) {
        const hipError_t q = hipEventQuery(e.ev);
        if (q == hipErrorNotReady) return MG_OK;
        if (q != hipSuccess) return set_error(MG_ERR_DEVICE, "engine event: %s", hipGetErrorString(q));
        e.inflight = false;
        const double t = s.now();
        for (Query *r : e.retiring) {
            if (e.kind != MG_Q_PPR) {
                r->te = e.b.h_sc->te[r->slot];
                r->nr = e.b.h_sc->nr[r->slot];
            }
            e.slot_q[r->slot] = nullptr;
            r->slot = -1;
            finish_query(s, r, t);
        }
        e.retiring.clear();
        if (e.kind != MG_Q_PPR && e.b.h_ctl->changed & (1 << 30))
            return set_error(MG_ERR_OVERFLOW, "serving engine task queue overflow");
        // slots whose search ended
        for (int sl = 0; sl < e.S; ++sl) {
            Query *r = e.slot_q[sl];
            if (!r) continue;
            bool fin;
            if (e.kind == MG_Q_PPR) {
                fin = r->iters_left <= 0;
            } else {
                const bool live = (e.b.h_sc->live[sl >> 5] >> (sl & 31)) & 1;
                fin = !live;
                if (fin && e.kind == MG_Q_SSSP) {
                    const long long cb = e.b.h_ctl->thr / e.delta - 1;
                    fin = (long long)e.b.h_sc->far_hi[sl] <= cb || e.b.h_ctl->done;
                }
            }
            if (fin) finished.push_back(sl);
        }
    }
    const bool has_waiting = !e.waiting.empty();
    if (finished.empty() && !has_waiting && e.running == 0) {
        if (e.idle_since == 0) e.idle_since = s.now();
        return MG_OK;
    }
    e.idle_since = 0;
    int rc;
    if (!e.allocated) {
        rc = engine_alloc(s, e);
        if (rc == MG_ERR_OOM && !e.allocated) {
            clear_error();
            return MG_OK;  // queries wait for memory
        }
        if (rc) return rc;
    }
    // 1. retire finished slots
    SlotList fl{};
    uint32_t fmask[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int sl : finished) {
        fl.slot[fl.n++] = sl;
        fmask[sl >> 5] |= 1u << (sl & 31);
    }
    if (fl.n) {
        if (e.kind == MG_Q_SSSP && (rc = sssp_batch_extract(e.b, e.g, e.st, fl))) return rc;
        if (e.kind == MG_Q_PPR) {
            if ((rc = ppr_extract_rows(e.p, e.st, fl))) return rc;
        } else if ((rc = batch_row_stats(e.b, e.g, e.st, fl, e.kind == MG_Q_BFS ? -1 : INT_MAX))) {
            return rc;
        }
        for (int sl : finished) {
            Query *r = e.slot_q[sl];
            const size_t bytes = 4 * (size_t)e.g->n;
            const void *row = e.kind == MG_Q_PPR ? (const void *)(e.p.rows + (size_t)sl * e.g->n)
                                                 : (const void *)(e.b.rows + (size_t)sl * e.g->n);
            if ((rc = e.d_perm ? deliver_twin_row(r, e.d_perm, e.h_perm, row, e.g->n, e.st)
                               : deliver_row(r, row, bytes, e.st)))
                return rc;
            if (e.kind == MG_Q_PPR) {
                r->te = e.g->m * (int64_t)(s.cfg.ppr_iters);
                r->nr = e.g->n;
            }
            e.retiring.push_back(r);
            e.running--;
        }
        if (e.kind != MG_Q_PPR && (rc = batch_release_slots(e.b, e.st, fmask))) return rc;
    }
    // 2. admit waiting queries into free slots
    SeedList seeds{};
    std::vector<bool> busy(e.S, false);
    for (int sl = 0; sl < e.S; ++sl) busy[sl] = e.slot_q[sl] != nullptr;
    const bool admit = e.kind != MG_Q_PPR || ppr_admits(s, e);
    for (int sl = 0; sl < e.S && admit && !e.waiting.empty(); ++sl) {
        if (busy[sl]) continue;
        Query *r = e.waiting.front();
        e.waiting.pop_front();
        r->slot = sl;
        r->t_admit = s.now();
        r->iters_left = s.cfg.ppr_iters;
        e.slot_q[sl] = r;
        busy[sl] = true;
        seeds.slot[seeds.n] = sl;
        seeds.src[seeds.n++] = e.h_perm ? e.h_perm[r->src] : r->src;
        e.running++;
    }
    if (e.kind == MG_Q_BFS) {
        if ((rc = bfs_batch_seed(e.b, e.g, e.st, seeds))) return rc;
    } else if (e.kind == MG_Q_SSSP) {
        if ((rc = sssp_batch_seed(e.b, e.g, e.st, seeds, e.delta, false, true))) return rc;
    } else if (seeds.n || fl.n) {
        if ((rc = ppr_seed(e.p, e.st, seeds, fl))) return rc;
    }
    // 3. a chunk of steps, then the status copy
    if (e.running > 0) {
        if (e.kind == MG_Q_PPR) {
            int it = 4;
            for (Query *r : e.slot_q)
                if (r) it = std::min(it, r->iters_left);
            it = std::max(it, 1);
            if ((rc = ppr_iterate(e.p, e.g, e.st, s.cfg.ppr_teleport, it))) return rc;
            for (Query *r : e.slot_q)
                if (r) r->iters_left -= it;
        } else if (e.chunk) {
            MG_HIP(hipGraphLaunch(e.chunk, e.st));
        } else {
            for (int i = 0; i < engine_steps(e); ++i)
                if ((rc = e.kind == MG_Q_BFS ? bfs_batch_step(e.b, e.g, e.st) : sssp_batch_step(e.b, e.g, e.st, e.delta)))
                    return rc;
            if ((rc = batch_live(e.b, e.st))) return rc;
        }
    }
    if (e.kind != MG_Q_PPR) {
        if ((rc = copy_out(e.h_ctl_dev, e.b.ctl, sizeof(Ctl), e.st))) return rc;
        if ((rc = copy_out(e.h_sc_dev, e.b.sc, sizeof(SlotCtl), e.st))) return rc;
    }
    MG_HIP(hipEventRecord(e.ev, e.st));
    e.inflight = true;
    *queued = true;
    s.stats.rounds++;
    return MG_OK;
}

// A plain engine's worker: takes its engine's queries in arrival order and runs each
// as an ordinary search on its own view of the registered graph (the graph's arrays
// and twin, the view's dedicated stream, workspace and lock), without the server lock
// held; PLAIN_WORKERS workers per engine overlap their searches.  A lightly loaded batched engine pays nearly a full round of device work
// per step whatever its slot count; one whole-GPU search per query costs the search
// alone (R-MAT scale 22 SSSP: about 1.3 ms).
constexpr int PLAIN_WORKERS = 2;  // searches of one plain engine in flight together

static void plain_worker(Server *sp, Engine *ep, mg_graph *view) {
    Server &s = *sp;
    Engine &e = *ep;
    (void)hipSetDevice(s.device);
    uint64_t seen = 0;
    for (;;) {
        {
            std::unique_lock<std::mutex> wl(e.wmu);
            e.wcv.wait(wl, [&] { return e.wake_gen != seen || e.quit; });
            if (e.quit) return;
            seen = e.wake_gen;
        }
        for (;;) {  // every query queued so far
            std::unique_lock<std::mutex> lk = fair_lock(s);
            if (s.stop || e.waiting.empty()) break;
            Query *r = e.waiting.front();
            e.waiting.pop_front();
            r->t_admit = s.now();
            e.running++;
            lk.unlock();
            mg_options o;
            mg_options_default(&o);
            o.out_on_device = r->out_on_device ? 1 : 0;
            mg_stats st;
            int32_t *out = (int32_t *)r->out;
            const int rc =
                e.kind == MG_Q_BFS ? mg_bfs(view, r->src, out, &o, &st) : mg_sssp(view, r->src, out, &o, &st);
            lk.lock();
            if (rc) {
                r->status = rc;
                s.error = rc;
            } else {
                r->te = st.edges_traversed;
                r->nr = st.vertices_reached;
            }
            e.running--;
            finish_query(s, r, s.now());
            s.stats.rounds++;
        }
    }
}

static void dispatcher(Server *sp) {
    Server &s = *sp;
    (void)hipSetDevice(s.device);
    std::unique_lock<std::mutex> lk(s.mu);
    while (!s.stop) {
        bool any = false, pending = false;
        const double t_pass = s.now();
        for (auto &ep : s.engines) {
            if (!ep) continue;
            bool q = false;
            const int rc = engine_round(s, *ep, &q);
            if (rc) {
                // a device error fails the engine's queries; the server keeps serving others
                s.error = rc;
                for (Query *r : ep->waiting) {
                    r->status = rc;
                    finish_query(s, r, s.now());
                }
                ep->waiting.clear();
                for (Query *&r : ep->slot_q)
                    if (r) {
                        r->status = rc;
                        finish_query(s, r, s.now());
                        r = nullptr;
                    }
                for (ThinSlot &ts : ep->tslots) {
                    for (Query *r : ts.qs) {
                        r->status = rc;
                        finish_query(s, r, s.now());
                    }
                    ts.qs.clear();
                }
                for (Query *r : ep->paused) {
                    r->status = rc;
                    finish_query(s, r, s.now());
                }
                ep->paused.clear();
                ep->running = 0;
            }
            any |= q;
            pending |= !ep->plain && (ep->inflight || ep->running > 0 || !ep->waiting.empty());
        }
        s.stats.dispatch_busy_ms += s.now() - t_pass;
        s.stats.dispatch_passes++;
        if (!any) {
            // nothing queued this pass: wait briefly for events or new queries
            s.cv_work.wait_for(lk, std::chrono::microseconds(pending ? 20 : 2000));
        } else if (s.want.load() > 0) {
            // others wait for the lock: let them in before the next pass (bounded)
            lk.unlock();
            for (int i = 0; i < 256 && s.want.load() > 0; ++i) std::this_thread::yield();
            lk.lock();
        }
    }
}

}  // namespace mg

using namespace mg;

struct mg_server : mg::Server {};

extern "C" {

void mg_server_config_default(mg_server_config *c) {
    std::memset(c, 0, sizeof *c);
    c->hbm_budget = 64LL << 30;
    c->max_slots = MG_BATCH_MAX;
    c->ppr_iters = 20;
    c->ppr_teleport = 0.15f;
    c->steps_per_round = STEPS_PER_GRAPH;
    c->sssp_delta = 0;
    c->plain_kinds = 1 << MG_Q_SSSP;
}

int mg_server_create(const mg_server_config *cfg, mg_server **out) {
    clear_error();
    if (!out) return set_error(MG_ERR_INVALID_ARG, "null output pointer");
    *out = nullptr;
    mg_server *s = new (std::nothrow) mg_server();
    if (!s) return set_error(MG_ERR_OOM, "host allocation failed");
    if (cfg) s->cfg = *cfg;
    else mg_server_config_default(&s->cfg);
    if (s->cfg.hbm_budget <= 0 || s->cfg.max_slots < 1 || s->cfg.max_slots > MG_BATCH_MAX || s->cfg.ppr_iters < 1 ||
        s->cfg.steps_per_round < 1 || s->cfg.steps_per_round > 64 || s->cfg.thin_streams < 0 || s->cfg.thin_streams > 8 ||
        s->cfg.thin_slice_us < -1 || s->cfg.ppr_gate_us < -1 || s->cfg.ppr_batch_min < 0 ||
        s->cfg.ppr_batch_min > PPR_SLOTS || (s->cfg.plain_kinds & ~3)) {
        delete s;
        return set_error(MG_ERR_INVALID_ARG, "invalid server configuration");
    }
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete s;
        return set_error(MG_ERR_NO_DEVICE, "no HIP device");
    }
    if (s->cfg.thin_slice_us >= 0) {
        int khz = 0;  // the wall clock that wall_clock64() reads in kernels
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->device) != hipSuccess || khz <= 0) {
            delete s;
            return set_error(MG_ERR_DEVICE, "device wall-clock rate unavailable");
        }
        const long long us = s->cfg.thin_slice_us > 0 ? s->cfg.thin_slice_us : 5000;
        s->thin_slice = std::max(1LL, us * khz / 1000);
    }
    s->th = std::thread(dispatcher, static_cast<Server *>(s));
    *out = s;
    return MG_OK;
}

int mg_server_add_graph(mg_server *s, mg_graph *g, int32_t *graph_id) {
    clear_error();
    if (!s || !g || !graph_id) return set_error(MG_ERR_INVALID_ARG, "null argument");
    if (g->device != s->device) return set_error(MG_ERR_INVALID_ARG, "graph lives on another device");
    // High-diameter graphs serve BFS / SSSP as thin searches: decided by the levels of
    // a probe BFS from the first vertex with out-edges (before the server lock: the
    // probe is an ordinary search on the graph)
    bool thin = false;
    const int thin_levels = s->cfg.thin_levels == 0 ? 256 : s->cfg.thin_levels;
    if (thin_levels > 0 && g->n > 0) {
        int probe = -1;
        std::vector<int32_t> ro(2);
        for (int v = 0; v < g->n && probe < 0; v += std::max(1, g->n / 64)) {
            if (copy_on(g->stream, ro.data(), g->row_off + v, 8, hipMemcpyDeviceToHost) != hipSuccess)
                return set_error(MG_ERR_DEVICE, "probe: offsets copy failed");
            if (ro[1] > ro[0]) probe = v;
        }
        if (probe >= 0) {
            mg_options po;
            mg_options_default(&po);
            po.out_on_device = 1;
            mg_stats st;
            const int rc = mg_bfs(g, probe, nullptr, &po, &st);
            if (rc) return rc;
            thin = st.iterations >= thin_levels;
        }
    }
    std::unique_lock<std::mutex> lk = fair_lock(*s);
    const int id = (int)s->graphs.size();
    s->graphs.push_back(g);
    for (int k = 0; k < 3; ++k) {
        auto e = std::make_unique<Engine>();
        e->srv = s;
        e->kind = k;
        e->thin = thin && k != MG_Q_PPR;
        e->plain = !e->thin && k != MG_Q_PPR && ((s->cfg.plain_kinds >> k) & 1);
        // batched engines search the twin (sources renamed on admission, rows mapped
        // back on delivery); a thin search writes its row straight into the output, so
        // thin engines keep the caller's numbering; plain engines call the graph's own
        // searches, which use its twin themselves
        e->g = g->twin && !e->thin && !e->plain ? g->twin : g;
        if (e->g != g) {
            e->h_perm = g->h_perm.data();
            e->d_perm = g->twin->perm;
        }
        if (e->plain) {
            // search views of the graph on dedicated streams: their searches never queue
            // behind another engine's kernels, nor behind the caller's own calls on g.
            // Views that do not fit the HBM budget leave the engine batched.
            int64_t total = 0;
            for (int w = 0; w < PLAIN_WORKERS; ++w) {
                mg_graph *v = nullptr;
                int rc = graph_view_create(g, &v);
                int64_t vb = 0;
                if (rc == MG_OK) rc = mg_graph_device_bytes(v, &vb);
                if (rc) {
                    if (v) mg_graph_destroy(v);
                    for (mg_graph *x : e->views) mg_graph_destroy(x);
                    return rc;
                }
                total += vb;
                e->views.push_back(v);
            }
            if (s->bytes_used + total <= s->cfg.hbm_budget) {
                s->bytes_used += total;
                for (mg_graph *v : e->views) e->workers.emplace_back(plain_worker, static_cast<Server *>(s), e.get(), v);
                s->engines.push_back(std::move(e));
                continue;
            }
            for (mg_graph *x : e->views) mg_graph_destroy(x);
            e->views.clear();
            e->plain = false;
            e->g = g->twin ? g->twin : g;
            if (e->g != g) {
                e->h_perm = g->h_perm.data();
                e->d_perm = g->twin->perm;
            }
        }
        if (e->thin) {
            // every launch slot on a dedicated stream: a thin launch runs for up to a
            // slice, and a pooled stream would put it ahead of other streams' work
            e->tslots.resize(thin_streams(*s));
            for (ThinSlot &ts : e->tslots)
                if (dedicated_stream(s->device, &ts.st) != hipSuccess ||
                    hipEventCreateWithFlags(&ts.ev, hipEventDisableTiming) != hipSuccess)
                    return set_error(MG_ERR_DEVICE, "stream/event creation failed");
            s->engines.push_back(std::move(e));
            continue;
        }
        if (hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&e->ev, hipEventDisableTiming) != hipSuccess)
            return set_error(MG_ERR_DEVICE, "stream/event creation failed");
        s->engines.push_back(std::move(e));
    }
    *graph_id = id;
    return MG_OK;
}

int mg_server_submit(mg_server *s, int32_t graph_id, int32_t kind, int32_t src, void *out, int32_t out_on_device,
                     int64_t *query_id) {
    clear_error();
    if (!s || !query_id) return set_error(MG_ERR_INVALID_ARG, "null argument");
    std::unique_lock<std::mutex> lk = fair_lock(*s);
    if (graph_id < 0 || graph_id >= (int)s->graphs.size()) return set_error(MG_ERR_INVALID_ARG, "unknown graph %d", graph_id);
    if (kind < MG_Q_BFS || kind > MG_Q_PPR) return set_error(MG_ERR_INVALID_ARG, "unknown query kind %d", kind);
    mg_graph *g = s->graphs[graph_id];
    if (src < 0 || src >= g->n) return set_error(MG_ERR_INVALID_ARG, "source %d outside [0,%d)", src, g->n);
    if (kind == MG_Q_SSSP && (!g->w || g->min_weight < 0))
        return set_error(MG_ERR_UNSUPPORTED, "SSSP needs non-negative edge weights");
    if (kind == MG_Q_PPR && !(g->flags & MG_GRAPH_SYMMETRIC) && !g->col_off)
        return set_error(MG_ERR_UNSUPPORTED, "personalized PageRank needs MG_GRAPH_SYMMETRIC or MG_GRAPH_BUILD_CSC");
    auto q = std::make_unique<Query>();
    q->id = s->next_id++;
    q->graph = graph_id;
    q->kind = kind;
    q->src = src;
    q->out = out;
    q->out_on_device = out_on_device != 0;
    q->out_dev = device_view(out, q->out_on_device);
    q->t_submit = s->now();
    Engine &e = *s->engines[(size_t)graph_id * 3 + kind];
    e.waiting.push_back(q.get());
    *query_id = q->id;
    s->queries.emplace(q->id, std::move(q));
    s->stats.submitted++;
    lk.unlock();
    if (e.plain) {
        {
            std::lock_guard<std::mutex> g(e.wmu);
            e.wake_gen++;
        }
        e.wcv.notify_all();
    } else {
        s->cv_work.notify_one();
    }
    return MG_OK;
}

static void fill_info(const Query &q, mg_query_info *info) {
    if (!info) return;
    std::memset(info, 0, sizeof *info);
    info->status = q.status;
    info->kind = q.kind;
    info->graph = q.graph;
    info->src = q.src;
    info->edges_traversed = q.te;
    info->vertices_reached = q.nr;
    info->t_submit_ms = q.t_submit;
    info->t_admit_ms = q.t_admit;
    info->t_done_ms = q.t_done;
}

int mg_server_wait(mg_server *s, int64_t query_id, int32_t timeout_ms, mg_query_info *info) {
    clear_error();
    if (!s) return set_error(MG_ERR_INVALID_ARG, "null server");
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::milliseconds(std::max(timeout_ms, 0));
    std::unique_lock<std::mutex> lk = fair_lock(*s);
    auto it = s->queries.find(query_id);
    if (it == s->queries.end()) return set_error(MG_ERR_INVALID_ARG, "unknown query %lld", (long long)query_id);
    Query *q = it->second.get();
    while (!q->done) {
        // sleep on the completion counter, not on mu (read under mu: a query that
        // finishes after this read bumps it)
        const uint64_t seen = s->done_seq.load();
        lk.unlock();
        {
            std::unique_lock<std::mutex> dl(s->dmu);
            auto moved = [&] { return s->done_seq.load() != seen; };
            if (timeout_ms < 0) s->dcv.wait(dl, moved);
            else s->dcv.wait_until(dl, deadline, moved);
        }
        lk = fair_lock(*s);
        if (!q->done && timeout_ms >= 0 && std::chrono::steady_clock::now() >= deadline)
            return set_error(MG_ERR_TIMEOUT, "query %lld not done after %d ms", (long long)query_id, timeout_ms);
    }
    fill_info(*q, info);
    const int st = q->status;
    s->queries.erase(query_id);  // by key: the map may have rehashed while the lock was released
    if (st) return set_error(st, "query %lld failed", (long long)query_id);
    return MG_OK;
}

int mg_server_now_ms(mg_server *s, double *ms) {
    clear_error();
    if (!s || !ms) return set_error(MG_ERR_INVALID_ARG, "null argument");
    *ms = s->now();
    return MG_OK;
}

int mg_server_get_stats(mg_server *s, mg_server_stats *st) {
    clear_error();
    if (!s || !st) return set_error(MG_ERR_INVALID_ARG, "null argument");
    std::unique_lock<std::mutex> lk = fair_lock(*s);
    *st = s->stats;
    st->device_bytes = s->bytes_used;
    st->thin_engines = 0;
    for (const auto &e : s->engines)
        if (e && e->thin) st->thin_engines++;
    return MG_OK;
}

int mg_server_destroy(mg_server *s) {
    clear_error();
    if (!s) return MG_OK;
    {
        std::unique_lock<std::mutex> lk = fair_lock(*s);
        s->stop = true;
        s->cv_work.notify_all();
    }
    for (auto &e : s->engines)
        if (e && e->plain) {
            {
                std::lock_guard<std::mutex> g(e->wmu);
                e->quit = true;
            }
            e->wcv.notify_all();
        }
    if (s->th.joinable()) s->th.join();
    for (auto &e : s->engines)
        if (e)
            for (std::thread &t : e->workers)
                if (t.joinable()) t.join();
    for (auto &e : s->engines)
        if (e)
            for (mg_graph *v : e->views) mg_graph_destroy(v);
    (void)hipSetDevice(s->device);
    for (auto &e : s->engines) {
        if (!e) continue;
        if (e->st) (void)hipStreamSynchronize(e->st);
        for (ThinSlot &ts : e->tslots)
            if (ts.st) (void)hipStreamSynchronize(ts.st);
        engine_free(*s, *e);
        if (e->ev) (void)hipEventDestroy(e->ev);
        if (e->st) (void)hipStreamDestroy(e->st);
        for (ThinSlot &ts : e->tslots) {
            if (ts.ev) (void)hipEventDestroy(ts.ev);
            if (ts.st) (void)hipStreamDestroy(ts.st);
        }
    }
    delete s;
    return MG_OK;
}

void *mg_host_alloc(size_t bytes) {
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

void mg_host_free(void *p) {
    if (p) (void)hipHostFree(p);
}

}  // extern "C"
