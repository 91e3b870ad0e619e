// enactor.h -- generic GAS enactor for MI355X (the role of
// GASengine/enactor_vertex_centric.cuh, EnactIterativeSearch at :2398-2966).
//
// Each iteration runs, as the reference does (enactor:2526-2872):
//   gather (over in/out/all edges)  ->  apply  ->  post_apply  ->  expand + contract
// with the reference's mode enums deciding which steps run.  What changes is how:
//   * frontier = bitmap; each iteration starts by compacting it into a sorted vertex
//     queue (count / scan / write, no atomics), so frontier order is deterministic;
//   * expand and contract are fused: expand_edge's output vertex goes straight
//     through contract and into the next-frontier bitmap (test-and-set dedupe), so
//     no O(m) edge frontier is materialised (the reference keeps three queues of
//     max_queue_sizing * m entries, csr_problem.cuh:627-827);
//   * lists of up to LANE_T edges are walked by one lane, longer ones are split into
//     TASK-edge tasks walked by a wave (the reference's CTA/warp/scan-gather split,
//     expand_atomic/cta.cuh:267-758, without its barriers);
//   * gather folds with gather_sum per vertex in a fixed order (lane-sequential,
//     or lane-strided then an xor butterfly), so results are deterministic;
//   * the loop runs on the device: kernels read the iteration state from device
//     memory and the host checks for termination only every few iterations (the
//     reference copies the frontier size back every iteration, enactor:2866).
// Vertices whose expand_vertex is false are not expanded (the reference's
// two-phase expand rule, enactor:360-389).  Edge ids passed to expand_edge and
// gather_edge are CSR edge ids also for in-edge (CSC) slots.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <type_traits>

#include "config.h"
#include "csr_problem.h"
#include "gas.h"

namespace GASengine {

namespace gas {

constexpr int BLOCK = 256;
constexpr int LANE_T = 32;
constexpr int TASK = 256;
constexpr int MAXG = 2048;

struct Ctl {
    int level;
    int done;
    int nf;          // frontier size of the current iteration
    int iter_limit;
    int nwide;       // gather: vertices handed to a wave
    int ntasks;      // expand: task records
    int bad_vertex;  // expand_edge produced a vertex id outside [0, n)
    int pad;
    long long edges;  // edges expanded
};

struct Task {
    int v, beg, end, view;
};

// Views of the graph: [0] out-edges (CSR), [1] in-edges (CSC or the CSR when symmetric).
struct Views {
    const int *off[2];
    const int *adj[2];
    const int *eid[2];  // CSR edge id per slot (nullptr: identity)
};

template <class Program>
struct Args {
    typedef typename Program::GatherType GatherType;
    Ctl *c;
    Views views;
    int n;
    int nwords;
    uint32_t *bm[2];
    int *q;
    int *counts;
    int *wide;
    Task *tasks;
    long long task_cap;
    char *changed;
    GatherType *gather_tmp;
    typename Program::VertexType vl;
    typename Program::EdgeType el;
    int gather_views;  // bit 0: out-edges, bit 1: in-edges
    int expand_views;
};

__device__ __forceinline__ int lane() { return __lane_id(); }

template <class T>
__device__ __forceinline__ T shfl_xor_any(T x, int m) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "GatherType must be a 4- or 8-byte type");
    if constexpr (sizeof(T) == 4) {
        int b;
        memcpy(&b, &x, 4);
        b = __shfl_xor(b, m);
        memcpy(&x, &b, 4);
    } else {
        long long b;
        memcpy(&b, &x, 8);
        b = __shfl_xor(b, m);
        memcpy(&x, &b, 8);
    }
    return x;
}

template <int B>
__device__ int block_scan_excl(int x, int *total) {
    __shared__ int s_w[B / 64];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o);
        if (l >= o) inc += y;
    }
    __syncthreads();
    if (l == 63) s_w[w] = inc;
    __syncthreads();
    int off = 0, tot = 0;
    for (int i = 0; i < B / 64; ++i) {
        if (i < w) off += s_w[i];
        tot += s_w[i];
    }
    *total = tot;
    return off + inc - x;
}

// ---- frontier compaction: bitmap -> sorted queue.  Workgroup b owns words
// [b * BLOCK * wpt, (b + 1) * BLOCK * wpt), thread t the wpt words from
// (b * BLOCK + t) * wpt.
template <class Program>
__global__ void __launch_bounds__(BLOCK) k_count(Args<Program> a, int wpt) {
    if (a.c->done) return;
    const int L = a.c->level;
    const uint32_t *bm = a.bm[L & 1];
    uint32_t *nx = a.bm[(L + 1) & 1];
    const int w0 = (blockIdx.x * BLOCK + threadIdx.x) * wpt;
    int cnt = 0;
    for (int w = w0; w < w0 + wpt && w < a.nwords; ++w) {
        cnt += __popc(bm[w]);
        nx[w] = 0;  // the next frontier starts empty
    }
    int tot;
    block_scan_excl<BLOCK>(cnt, &tot);
    if (threadIdx.x == 0) a.counts[blockIdx.x] = tot;
}

template <class Program>
__global__ void __launch_bounds__(1024) k_scan(Args<Program> a, int nblocks) {
    Ctl *c = a.c;
    if (c->done) return;
    __shared__ int s_run;
    if (threadIdx.x == 0) s_run = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nblocks; b0 += 1024) {
        const int b = b0 + threadIdx.x;
        const int x = b < nblocks ? a.counts[b] : 0;
        int tot;
        const int ex = block_scan_excl<1024>(x, &tot);
        if (b < nblocks) a.counts[MAXG + b] = s_run + ex;
        __syncthreads();
        if (threadIdx.x == 0) s_run += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        c->nf = s_run;
        c->nwide = 0;
        c->ntasks = 0;
        if (s_run == 0 || c->level >= c->iter_limit) c->done = 1;
    }
}

template <class Program>
__global__ void __launch_bounds__(BLOCK) k_write(Args<Program> a, int wpt) {
    if (a.c->done) return;
    const uint32_t *bm = a.bm[a.c->level & 1];
    const int w0 = (blockIdx.x * BLOCK + threadIdx.x) * wpt;
    int cnt = 0;
    for (int w = w0; w < w0 + wpt && w < a.nwords; ++w) cnt += __popc(bm[w]);
    int tot;
    int k = a.counts[MAXG + blockIdx.x] + block_scan_excl<BLOCK>(cnt, &tot);
    for (int w = w0; w < w0 + wpt && w < a.nwords; ++w) {
        uint32_t word = bm[w];
        while (word) {
            const int b = __ffs(word) - 1;
            a.q[k++] = w * 32 + b;
            word &= word - 1;
        }
    }
}

// ---- gather: fold gather_edge over the selected views with gather_sum
template <class Program>
__device__ __forceinline__ int gather_degree(const Args<Program> &a, int v) {
    int d = 0;
    for (int vw = 0; vw < 2; ++vw)
        if (a.gather_views & (1 << vw)) d += a.views.off[vw][v + 1] - a.views.off[vw][v];
    return d;
}

template <class Program>
__device__ __forceinline__ void gather_fold(Args<Program> &a, int v, int first, int stride, bool &has,
                                            typename Program::GatherType &acc) {
    typename Program::gather_edge ge;
    typename Program::gather_sum gs;
    int k0 = 0;  // position of this view's first edge in the concatenated list
    for (int vw = 0; vw < 2; ++vw) {
        if (!(a.gather_views & (1 << vw))) continue;
        const int beg = a.views.off[vw][v], end = a.views.off[vw][v + 1];
        // first local index j >= 0 with k0 + j = first (mod stride)
        int j = first - k0;
        if (j < 0) j = ((-j) % stride) ? stride - ((-j) % stride) : 0;
        for (int e = beg + j; e < end; e += stride) {
            const int eid = a.views.eid[vw] ? a.views.eid[vw][e] : e;
            typename Program::GatherType val;
            ge(v, eid, a.views.adj[vw][e], a.vl, a.el, val);
            acc = has ? gs(acc, val) : val;
            has = true;
        }
        k0 += end - beg;
    }
}

template <class Program>
__global__ void __launch_bounds__(BLOCK) k_gather_lanes(Args<Program> a) {
    if (a.c->done) return;
    __shared__ int s_base;
    const int nf = a.c->nf;
    for (int i0 = blockIdx.x * BLOCK; i0 < nf; i0 += gridDim.x * BLOCK) {
        const int i = i0 + threadIdx.x;
        const int v = i < nf ? a.q[i] : -1;
        const int d = v >= 0 ? gather_degree(a, v) : 0;
        const bool wide = d > LANE_T;
        int tot;
        const int ex = block_scan_excl<BLOCK>(wide ? 1 : 0, &tot);
        if (tot) {
            if (threadIdx.x == 0) s_base = atomicAdd(&a.c->nwide, tot);
            __syncthreads();
            if (wide) a.wide[s_base + ex] = v;
        }
        if (v >= 0 && !wide) {
            bool has = false;
            typename Program::GatherType acc = Program::INIT_VALUE;
            gather_fold(a, v, 0, 1, has, acc);
            a.gather_tmp[v] = has ? acc : (typename Program::GatherType)Program::INIT_VALUE;
        }
        __syncthreads();
    }
}

template <class Program>
__global__ void __launch_bounds__(BLOCK) k_gather_wide(Args<Program> a) {
    if (a.c->done) return;
    typedef typename Program::GatherType G;
    typename Program::gather_sum gs;
    const int nw = a.c->nwide;
    const int wave = (blockIdx.x * BLOCK + threadIdx.x) / 64, nwaves = gridDim.x * BLOCK / 64;
    for (int t = wave; t < nw; t += nwaves) {
        const int v = a.wide[t];
        bool has = false;
        G acc = (G)Program::INIT_VALUE;
        gather_fold(a, v, lane(), 64, has, acc);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {  // xor butterfly: the same tree on every run
            const G other = shfl_xor_any(acc, o);
            const bool oh = __shfl_xor((int)has, o) != 0;
            if (has && oh) acc = (lane() & o) ? gs(other, acc) : gs(acc, other);
            else if (oh) acc = other;
            has = has || oh;
        }
        if (lane() == 0) a.gather_tmp[v] = has ? acc : (G)Program::INIT_VALUE;
    }
}

// ---- apply / post_apply over the frontier (or all vertices)
template <class Program>
__global__ void __launch_bounds__(BLOCK) k_apply(Args<Program> a, int all) {
    if (a.c->done) return;
    typename Program::apply ap;
    const int iter = a.c->level;
    const int cnt = all ? a.n : a.c->nf;
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < cnt; i += gridDim.x * BLOCK) {
        const int v = all ? i : a.q[i];
        ap(v, iter, a.gather_tmp[v], a.vl, a.el, a.changed[v]);
    }
}

template <class Program>
__global__ void __launch_bounds__(BLOCK) k_post_apply(Args<Program> a, int all) {
    if (a.c->done) return;
    typename Program::post_apply pa;
    const int cnt = all ? a.n : a.c->nf;
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < cnt; i += gridDim.x * BLOCK) {
        const int v = all ? i : a.q[i];
        pa(v, a.vl, a.el, a.gather_tmp);
    }
}

// ---- expand + contract, fused
template <class Program>
__device__ __forceinline__ void expand_one(Args<Program> &a, bool changed, int iter, int v, int vw, int e) {
    typename Program::expand_edge ee;
    typename Program::contract ct;
    const int nbr = a.views.adj[vw][e];
    const int eid = a.views.eid[vw] ? a.views.eid[vw][e] : e;
    int f = -1;
    typename Program::MiscType misc = typename Program::MiscType();
    ee(changed, iter, v, nbr, eid, a.vl, a.el, f, misc);
    if (f == -1) return;
    if (f < 0 || f >= a.n) {
        a.c->bad_vertex = 1;
        return;
    }
    ct(iter, f, a.vl, a.el, a.gather_tmp, misc);
    if (f == -1) return;
    uint32_t *nx = a.bm[(iter + 1) & 1];
    const uint32_t bit = 1u << (f & 31);
    if (!(nx[f >> 5] & bit)) atomicOr(&nx[f >> 5], bit);
}

template <class Program>
__global__ void __launch_bounds__(BLOCK) k_expand_lanes(Args<Program> a) {
    if (a.c->done) return;
    __shared__ int s_base;
    typename Program::expand_vertex xv;
    const int nf = a.c->nf, iter = a.c->level;
    for (int i0 = blockIdx.x * BLOCK; i0 < nf; i0 += gridDim.x * BLOCK) {
        const int i = i0 + threadIdx.x;
        const int v = i < nf ? a.q[i] : -1;
        const bool go = v >= 0 && xv(v, a.changed[v], a.vl, a.el);
        for (int vw = 0; vw < 2; ++vw) {
            if (!(a.expand_views & (1 << vw))) continue;
            int beg = 0, deg = 0;
            if (go) {
                beg = a.views.off[vw][v];
                deg = a.views.off[vw][v + 1] - beg;
            }
            const int nt = deg > LANE_T ? (deg + TASK - 1) / TASK : 0;
            int tot;
            const int tex = block_scan_excl<BLOCK>(nt, &tot);
            if (tot) {
                if (threadIdx.x == 0) s_base = atomicAdd(&a.c->ntasks, tot);
                __syncthreads();
                for (int k = 0; k < nt; ++k) {
                    const long long slot = (long long)s_base + tex + k;
                    if (slot < a.task_cap)
                        a.tasks[slot] = Task{v, beg + k * TASK, min(beg + (k + 1) * TASK, beg + deg), vw};
                    else
                        a.c->bad_vertex = 2;
                }
                __syncthreads();
            }
            if (nt == 0)
                for (int e = beg; e < beg + deg; ++e) expand_one(a, true, iter, v, vw, e);
        }
    }
}

template <class Program>
__global__ void __launch_bounds__(BLOCK) k_expand_tasks(Args<Program> a) {
    if (a.c->done) return;
    const int nt = (int)min((long long)a.c->ntasks, a.task_cap), iter = a.c->level;
    const int wave = (blockIdx.x * BLOCK + threadIdx.x) / 64, nwaves = gridDim.x * BLOCK / 64;
    for (int t = wave; t < nt; t += nwaves) {
        const Task r = a.tasks[t];
        for (int e = r.beg + lane(); e < r.end; e += 64) expand_one(a, true, iter, r.v, r.view, e);
    }
}

template <class Program>
__global__ void k_advance(Args<Program> a) {
    if (!a.c->done) a.c->level++;
}

template <class Program>
__global__ void k_seed(Args<Program> a, const int *keys, int nkeys, int iter_limit) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nkeys; i += gridDim.x * blockDim.x) {
        const int v = keys[i];
        if (v >= 0 && v < a.n) atomicOr(&a.bm[0][v >> 5], 1u << (v & 31));
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.c->level = 0;
        a.c->done = iter_limit <= 0;
        a.c->nf = 0;
        a.c->iter_limit = iter_limit;
        a.c->nwide = 0;
        a.c->ntasks = 0;
        a.c->bad_vertex = 0;
        a.c->edges = 0;
    }
}

}  // namespace gas

// The enactor: EnactorVertexCentric<CsrProblem, Program, INSTRUMENT>(cfg, DEBUG),
// then EnactIterativeSearch(problem, h_row_offsets, directed, num_srcs, srcs,
// iter_num, threshold) as in enactor_vertex_centric.cuh:2910-2966.
template <typename CsrProblemT, typename Program, bool INSTRUMENT>
class EnactorVertexCentric {
  public:
    typedef typename CsrProblemT::SizeT SizeT;
    typedef typename Program::GatherType GatherType;

    explicit EnactorVertexCentric(const mapgraph::Config &cfg = mapgraph::Config(), bool debug = false)
        : cfg_(cfg), debug_(debug) {}
    EnactorVertexCentric(const EnactorVertexCentric &) = delete;
    EnactorVertexCentric &operator=(const EnactorVertexCentric &) = delete;
    ~EnactorVertexCentric() { release(); }

    // Runs the program to convergence (empty frontier) or iter_num iterations.
    // h_row_offsets and threshold are accepted for signature compatibility (the
    // reference also ignores h_row_offsets, enactor:2401; threshold picked its
    // expand strategy, which the lane/task scheme makes unnecessary).
    hipError_t EnactIterativeSearch(CsrProblemT &problem, SizeT *h_row_offsets, int directed, int num_srcs,
                                    int *srcs, int iter_num, int threshold) {
        (void)h_row_offsets, (void)threshold;
        if (problem.graph_slices.size() != 1) return hipErrorInvalidValue;
        auto *s = problem.graph_slices[0];
        hipError_t e;
        if ((e = hipSetDevice(s->gpu)) != hipSuccess) return e;
        const int n = s->nodes;
        if ((e = reserve(n, s->edges)) != hipSuccess) return e;
        hipStream_t st = s->stream;

        gas::Args<Program> a{};
        a.c = ctl_;
        const bool have_in = s->symmetric || s->d_column_offsets;
        a.views.off[0] = s->d_row_offsets;
        a.views.adj[0] = s->d_column_indices;
        a.views.eid[0] = nullptr;
        if (s->symmetric) {
            a.views.off[1] = s->d_row_offsets;
            a.views.adj[1] = s->d_column_indices;
            a.views.eid[1] = nullptr;
        } else {
            a.views.off[1] = s->d_column_offsets;
            a.views.adj[1] = s->d_row_indices;
            a.views.eid[1] = s->d_edgeCSC_indices;
        }
        a.n = n;
        a.nwords = nwords_;
        a.bm[0] = bm_[0];
        a.bm[1] = bm_[1];
        a.q = q_;
        a.counts = counts_;
        a.wide = wide_;
        a.tasks = tasks_;
        a.task_cap = task_cap_;
        a.changed = s->d_changed;
        a.gather_tmp = gather_tmp_;
        // views per mode; directed == 0 adds the other direction as the reference
        // does (enactor:1889-1965 for expand, 1479-1487 for gather)
        const bool sym = s->symmetric;
        auto pick = [&](int out, int in) { return (out ? 1 : 0) | (in ? 2 : 0); };
        switch (Program::gatherOverEdges()) {
            case NO_GATHER_EDGES: a.gather_views = 0; break;
            case GATHER_IN_EDGES: a.gather_views = pick(directed == 0 && !sym, 1); break;
            case GATHER_OUT_EDGES: a.gather_views = pick(1, directed == 0 && !sym); break;
            default: a.gather_views = sym ? 1 : 3; break;
        }
        switch (Program::expandOverEdges()) {
            case NO_EXPAND_EDGES: a.expand_views = 0; break;
            case EXPAND_IN_EDGES: a.expand_views = pick(directed == 0 && !sym, 1); break;
            case EXPAND_OUT_EDGES: a.expand_views = pick(1, directed == 0 && !sym); break;
            default: a.expand_views = sym ? 1 : 3; break;
        }
        if (((a.gather_views | a.expand_views) & 2) && !have_in) return hipErrorInvalidValue;

        auto t0 = std::chrono::steady_clock::now();
        // gather_tmp = INIT_VALUE (enactor:2473-2478), changed = 0
        if ((e = mapgraph::fill(gather_tmp_, (GatherType)Program::INIT_VALUE, (size_t)n, st)) != hipSuccess) return e;
        if ((e = mapgraph::fill(s->d_changed, (char)0, (size_t)n, st)) != hipSuccess) return e;
        if ((e = mapgraph::fill(bm_[0], 0u, (size_t)nwords_, st)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
        const int nsrc = Program::srcVertex() == ALL ? n : num_srcs;
        Program::Initialize(directed, n, (int)s->edges, nsrc, srcs, s->d_row_offsets, s->d_column_indices,
                            s->d_column_offsets, s->d_row_indices, s->d_edge_values, s->vertex_list, s->edge_list,
                            s->frontier_queues.d_keys, s->frontier_queues.d_values);
        if ((e = hipDeviceSynchronize()) != hipSuccess) return e;
        a.vl = s->vertex_list;
        a.el = s->edge_list;

        hipEvent_t ev0, ev1;
        (void)hipEventCreate(&ev0);
        (void)hipEventCreate(&ev1);
        (void)hipEventRecord(ev0, st);
        // compaction: wpt words per thread so that Gw workgroups cover the bitmap
        const int wpt = std::max(1, (nwords_ + gas::MAXG * gas::BLOCK - 1) / (gas::MAXG * gas::BLOCK));
        const unsigned Gw =
            (unsigned)std::max(1, std::min(gas::MAXG, (nwords_ + gas::BLOCK * wpt - 1) / (gas::BLOCK * wpt)));
        const unsigned G = (unsigned)std::max(1, std::min(gas::MAXG, (n + gas::BLOCK - 1) / gas::BLOCK));
        hipLaunchKernelGGL(gas::k_seed<Program>, dim3(G), dim3(gas::BLOCK), 0, st, a,
                           (const int *)s->frontier_queues.d_keys[0], std::max(0, std::min(nsrc, n)), iter_num);
        const bool do_gather = Program::gatherOverEdges() != NO_GATHER_EDGES;
        const int apply_mode = Program::applyOverEdges(), post_mode = Program::postApplyOverEdges();
        auto level = [&]() {
            hipLaunchKernelGGL(gas::k_count<Program>, dim3(Gw), dim3(gas::BLOCK), 0, st, a, wpt);
            hipLaunchKernelGGL(gas::k_scan<Program>, dim3(1), dim3(1024), 0, st, a, (int)Gw);
            hipLaunchKernelGGL(gas::k_write<Program>, dim3(Gw), dim3(gas::BLOCK), 0, st, a, wpt);
            if (do_gather) {
                hipLaunchKernelGGL(gas::k_gather_lanes<Program>, dim3(G), dim3(gas::BLOCK), 0, st, a);
                hipLaunchKernelGGL(gas::k_gather_wide<Program>, dim3(G), dim3(gas::BLOCK), 0, st, a);
            }
            if (apply_mode != NO_APPLY_VERTICES)
                hipLaunchKernelGGL(gas::k_apply<Program>, dim3(G), dim3(gas::BLOCK), 0, st, a,
                                   (int)(apply_mode == APPLY_ALL));
            if (post_mode != NO_POST_APPLY_VERTICES)
                hipLaunchKernelGGL(gas::k_post_apply<Program>, dim3(G), dim3(gas::BLOCK), 0, st, a,
                                   (int)(post_mode == POST_APPLY_ALL));
            if (a.expand_views) {
                hipLaunchKernelGGL(gas::k_expand_lanes<Program>, dim3(G), dim3(gas::BLOCK), 0, st, a);
                hipLaunchKernelGGL(gas::k_expand_tasks<Program>, dim3(G), dim3(gas::BLOCK), 0, st, a);
            }
            hipLaunchKernelGGL(gas::k_advance<Program>, dim3(1), dim3(1), 0, st, a);
        };
        // poll every 1, 2, 4, ... 32 queued iterations
        long long queued = 0, k = 1;
        const long long max_levels = (long long)std::min(iter_num, n + 1) + 1;
        gas::Ctl h{};
        for (;;) {
            for (long long i = 0; i < k && queued <= max_levels; ++i, ++queued) level();
            if ((e = hipGetLastError()) != hipSuccess) return e;
            if ((e = hipMemcpyAsync(&h, ctl_, sizeof h, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
            if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
            if (debug_) fprintf(stderr, "Iteration: %d, frontier_size: %d\n", h.level, h.nf);
            if (h.done || queued > max_levels) break;
            if (k < 32) k *= 2;
        }
        (void)hipEventRecord(ev1, st);
        (void)hipEventSynchronize(ev1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ev0, ev1);
        (void)hipEventDestroy(ev0);
        (void)hipEventDestroy(ev1);
        kernel_ms_ = ms;
        wall_ms_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        iterations_ = h.level;
        if (h.bad_vertex) return hipErrorLaunchFailure;
        return hipGetLastError();
    }

    // "Kernel time took" / "Total iteration" of the reference (enactor:2874-2893).
    double kernel_ms() const { return kernel_ms_; }
    double wall_ms() const { return wall_ms_; }
    int iterations() const { return iterations_; }

  private:
    hipError_t reserve(int n, long long m) {
        if (n <= cap_n_ && m <= cap_m_) return hipSuccess;
        release();
        cap_n_ = n;
        cap_m_ = m;
        nwords_ = ((n + 63) / 64) * 2;
        task_cap_ = 2 * (m / gas::TASK + m / (gas::LANE_T + 1)) + 64;
        hipError_t e;
        const size_t nn = n > 0 ? (size_t)n : 1;
        if ((e = hipMalloc((void **)&ctl_, sizeof(gas::Ctl))) != hipSuccess ||
            (e = hipMalloc((void **)&bm_[0], 4 * (size_t)nwords_ + 4)) != hipSuccess ||
            (e = hipMalloc((void **)&bm_[1], 4 * (size_t)nwords_ + 4)) != hipSuccess ||
            (e = hipMalloc((void **)&q_, 4 * nn)) != hipSuccess ||
            (e = hipMalloc((void **)&wide_, 4 * nn)) != hipSuccess ||
            (e = hipMalloc((void **)&counts_, 8 * gas::MAXG)) != hipSuccess ||
            (e = hipMalloc((void **)&tasks_, sizeof(gas::Task) * (size_t)task_cap_)) != hipSuccess ||
            (e = hipMalloc((void **)&gather_tmp_, sizeof(GatherType) * nn)) != hipSuccess)
            return e;
        return hipSuccess;
    }

    void release() {
        void *ps[] = {ctl_, bm_[0], bm_[1], q_, wide_, counts_, tasks_, gather_tmp_};
        for (void *p : ps)
            if (p) (void)hipFree(p);
        ctl_ = nullptr;
        bm_[0] = bm_[1] = nullptr;
        q_ = wide_ = counts_ = nullptr;
        tasks_ = nullptr;
        gather_tmp_ = nullptr;
        cap_n_ = -1;
        cap_m_ = -1;
    }

    mapgraph::Config cfg_;
    bool debug_;
    gas::Ctl *ctl_ = nullptr;
    uint32_t *bm_[2] = {nullptr, nullptr};
    int *q_ = nullptr, *wide_ = nullptr, *counts_ = nullptr;
    gas::Task *tasks_ = nullptr;
    GatherType *gather_tmp_ = nullptr;
    int nwords_ = 0;
    long long task_cap_ = 0;
    int cap_n_ = -1;
    long long cap_m_ = -1;
    double kernel_ms_ = 0, wall_ms_ = 0;
    int iterations_ = 0;
};

}  // namespace GASengine
