// Thin searches for serving: one workgroup runs one query's whole BFS or SSSP.
//
// Not in the reference (its drivers run one search at a time over the whole GPU,
// Algorithms/BFS/bfs.cu:402-417).  On high-diameter graphs (grids: thousands of
// levels of a few thousand vertices each) a search is a long chain of small levels;
// as batched whole-GPU levels each costs a launch and a queue slot, and in a serving
// mix those queue behind other engines' kernels.  Here a launch carries one
// workgroup per admitted query, and each workgroup runs its query to the end inside
// the launch, levels back to back with workgroup barriers, on its own lane of device
// state: the result row doubles as the depth / distance array, beside two frontier
// queues, a far list and two bitmaps.  Concurrent queries are concurrent workgroups
// (one CU each), so dozens of grid queries share the GPU with no per-level launch.
// A launch is time-sliced: a search that has run `slice` wall-clock ticks pauses at
// the next level boundary (its queues and bitmaps stay in its lane; four scalars are
// saved) and the server launches it again with the next admitted queries, so no
// launch holds its hardware queue (which other streams may share) for long, and new
// queries start within a slice.  A finished search copies its row straight into the
// query's output.
//
// Semantics are the plain searches': BFS depths (level-synchronous, a vertex takes
// the level of its first claim; unreached -1) and SSSP distances (delta-stepping
// with re-insertion, so label-correcting: the unique shortest distances, unreached
// INT_MAX).  Only this workgroup touches a lane during the launch, so loads and
// atomics take workgroup scope.
#include <climits>

#include "thin.h"
#include "frontier.h"

namespace mg {

constexpr int THIN_BLK = 1024;
constexpr int THIN_EW = 4;  // edge slots per thread in flight

__device__ __forceinline__ bool cas_wg(int *p, int expect, int val) {
    return __hip_atomic_compare_exchange_strong(p, &expect, val, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ int fetch_min_wg(int *p, int v) {
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Appends u (lanes with `want`) to q, one LDS counter reservation per wave.
__device__ __forceinline__ void thin_append(bool want, int u, int *q, int *s_cnt) {
    const uint64_t mask = wave_ballot(want);
    if (!mask) return;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    int base = 0;
    if (lane_id() == leader) base = atomicAdd(s_cnt, __popcll(mask));
    base = __shfl(base, leader);
    if (want) q[base + lanes_below(mask)] = u;
}

// Walks the out-edges of queue entries [t0, t0 + THIN_BLK) of q (nq entries):
// prefix sums of the lengths in LDS, each slot's owner found by binary search, so the
// threads stream the concatenated lists THIN_EW slots each in flight.
// visit(u, w, owner_index, ok) is called in lock-step by every thread.
template <bool WEIGHTED, class Load, class Visit>
__device__ __forceinline__ void thin_tile(const ThinArgs &a, const int *q, int nq, int t0, int *s_pre, int *s_beg,
                                          int *s_val, int *s_wave, Load load, Visit visit) {
    const int i = t0 + (int)threadIdx.x;
    int beg = 0, deg = 0, val = 0;
    if (i < nq) {
        const int v = ld_wg(&q[i]);
        beg = a.off[v];
        deg = a.off[v + 1] - beg;
        val = load(v, deg);
    }
    int total;
    const int ex = block_exclusive_scan<THIN_BLK>(deg, s_wave, &total);
    s_pre[threadIdx.x] = ex;
    s_beg[threadIdx.x] = beg;
    s_val[threadIdx.x] = val;
    __syncthreads();
    const int nown = min(THIN_BLK, nq - t0);
    for (int p0 = 0; p0 < total; p0 += THIN_EW * THIN_BLK) {
        int sl[THIN_EW], own[THIN_EW];
        bool ok[THIN_EW];
#pragma unroll
        for (int j = 0; j < THIN_EW; ++j) {
            int p = p0 + j * THIN_BLK + (int)threadIdx.x;
            ok[j] = p < total;
            p = ok[j] ? p : total - 1;  // past the end: a real slot, ignored
            int o = 0;                  // the last list whose prefix is <= p (a non-empty one)
            for (int st = 1 << (31 - __builtin_clz(nown)); st > 0; st >>= 1)
                if (o + st < nown && s_pre[o + st] <= p) o += st;
            own[j] = o;
            sl[j] = s_beg[o] + (p - s_pre[o]);
        }
        int u[THIN_EW], w[THIN_EW];
#pragma unroll
        for (int j = 0; j < THIN_EW; ++j) u[j] = a.adj[sl[j]];
#pragma unroll
        for (int j = 0; j < THIN_EW; ++j) w[j] = WEIGHTED ? a.w[sl[j]] : 0;
        visit(u, w, own, ok);
    }
    __syncthreads();  // the LDS arrays are rewritten by the next tile
}

// Sum of the out-degrees of the reached vertices (the reference TEPS count,
// bfs.cu:483-493) and their number, from the finished row.
__device__ void thin_row_stats(const ThinArgs &a, const int *row, int unreached, long long *out) {
    long long te = 0, nr = 0;
    for (int v = threadIdx.x; v < a.n; v += THIN_BLK)
        if (ld_wg(&row[v]) != unreached) {
            te += a.off[v + 1] - a.off[v];
            ++nr;
        }
    te = block_sum<THIN_BLK>(te);
    nr = block_sum<THIN_BLK>(nr);
    if (threadIdx.x == 0) {
        out[0] = te;
        out[1] = nr;
    }
}

// The finished row into the query's output (device memory or pinned host memory),
// 16 bytes a store where both sides allow it.
__device__ void thin_copy_row(const int *row, int n, void *out) {
    int *o = (int *)out;
    int v0 = 0;
    if ((((uintptr_t)o | (uintptr_t)row) & 15) == 0) {
        const uint4 *r4 = (const uint4 *)row;
        uint4 *o4 = (uint4 *)o;
        for (int i = threadIdx.x; i < n / 4; i += THIN_BLK) o4[i] = r4[i];
        v0 = n / 4 * 4;
    }
    for (int v = v0 + (int)threadIdx.x; v < n; v += THIN_BLK) o[v] = row[v];
}

// Slice check at a level boundary (thread 0 decides, the barrier after it publishes
// the decision): a search pauses once it has run at least one level in this launch
// and `slice` ticks have passed since the launch started.
__device__ __forceinline__ void thin_slice_check(const ThinBatch &jobs, long long t_start, int levels, int *s_stop) {
    if (threadIdx.x == 0)
        *s_stop = jobs.slice > 0 && levels > 0 && (long long)wall_clock64() - t_start >= jobs.slice;
}

// Finish: the row's counts and the finished flag, then the row into the output.
__device__ void thin_finish(const ThinArgs &a, const ThinBatch &jobs, const int *row, int unreached, int lane_no) {
    thin_row_stats(a, row, unreached, a.stats + THIN_STAT * lane_no);
    if (threadIdx.x == 0) a.stats[THIN_STAT * lane_no + 2] = 1;
    if (jobs.out[blockIdx.x]) thin_copy_row(row, a.n, jobs.out[blockIdx.x]);
}

__global__ __launch_bounds__(THIN_BLK) void thin_bfs(ThinArgs a, ThinBatch jobs) {
    __shared__ int s_pre[THIN_BLK], s_beg[THIN_BLK], s_val[THIN_BLK], s_wave[THIN_BLK / WAVE];
    __shared__ int s_n, s_cnt, s_stop;
    const long long t_start = wall_clock64();
    const int lane_no = jobs.lane[blockIdx.x], src = jobs.src[blockIdx.x];
    const size_t n = (size_t)a.n;
    int *row = a.rows + lane_no * n;
    int *q0 = a.q + lane_no * 2 * n, *q1 = q0 + n;
    long long *st = a.state + 4 * (size_t)lane_no;
    int par = 0, L = 0;
    if (src >= 0) {
        for (int v = threadIdx.x; v < a.n; v += THIN_BLK) row[v] = v == src ? 0 : -1;
        if (threadIdx.x == 0) {
            q0[0] = src;
            s_n = 1;
        }
    } else {  // resume: queue parity, queue size, level
        par = (int)st[0];
        L = (int)st[2];
        if (threadIdx.x == 0) s_n = (int)st[1];
    }
    int *qa = par ? q1 : q0, *qb = par ? q0 : q1;
    __syncthreads();
    for (int done = 0;; ++L, ++done) {
        const int nq = s_n;
        if (nq == 0) break;
        thin_slice_check(jobs, t_start, done, &s_stop);
        __syncthreads();  // every thread has read s_n; s_stop is set
        if (s_stop) {
            if (threadIdx.x == 0) {
                st[0] = par;
                st[1] = nq;
                st[2] = L;
                a.stats[THIN_STAT * lane_no + 2] = 0;
            }
            return;
        }
        if (threadIdx.x == 0) s_cnt = 0;
        for (int t0 = 0; t0 < nq; t0 += THIN_BLK)
            thin_tile<false>(a, qa, nq, t0, s_pre, s_beg, s_val, s_wave, [](int, int) { return 0; },
                             [&](const int(&u)[THIN_EW], const int(&)[THIN_EW], const int(&)[THIN_EW],
                                 const bool(&ok)[THIN_EW]) {
                                 int cur[THIN_EW];
#pragma unroll
                                 for (int j = 0; j < THIN_EW; ++j) cur[j] = ld_wg(&row[u[j]]);
#pragma unroll
                                 for (int j = 0; j < THIN_EW; ++j) {
                                     const bool newly = ok[j] && cur[j] == -1 && cas_wg(&row[u[j]], -1, L + 1);
                                     thin_append(newly, u[j], qb, &s_cnt);
                                 }
                             });
        __syncthreads();
        if (threadIdx.x == 0) s_n = s_cnt;
        int *t = qa;
        qa = qb;
        qb = t;
        par ^= 1;
        __syncthreads();
    }
    thin_finish(a, jobs, row, -1, lane_no);
}

// Delta-stepping: the near queue holds vertices improved below the bucket bound thr
// (de-duplicated by the `inq` bitmap); improvements at or above it go to the far list
// (de-duplicated by `farbm`).  When the near queue empties, thr moves to the bucket of
// the smallest waiting distance and the far entries below it join the near queue;
// entries that dropped below the old bound were relaxed as near vertices already.
__global__ __launch_bounds__(THIN_BLK) void thin_sssp(ThinArgs a, ThinBatch jobs) {
    __shared__ int s_pre[THIN_BLK], s_beg[THIN_BLK], s_val[THIN_BLK], s_wave[THIN_BLK / WAVE];
    __shared__ int s_n, s_cnt, s_nfar, s_kept, s_stop;
    __shared__ long long s_thr;
    const long long t_start = wall_clock64();
    const int lane_no = jobs.lane[blockIdx.x], src = jobs.src[blockIdx.x];
    const size_t n = (size_t)a.n;
    int *row = a.rows + lane_no * n;
    int *q0 = a.q + lane_no * 2 * n, *q1 = q0 + n;
    int *far = a.far + lane_no * n;
    uint32_t *inq = a.bits + lane_no * 2 * (size_t)a.nwords, *farbm = inq + a.nwords;
    long long *st = a.state + 4 * (size_t)lane_no;
    const long long delta = a.delta;
    int par = 0;
    if (src >= 0) {
        for (int v = threadIdx.x; v < a.n; v += THIN_BLK) row[v] = v == src ? 0 : INT_MAX;
        for (int i = threadIdx.x; i < 2 * a.nwords; i += THIN_BLK) inq[i] = 0;
        if (threadIdx.x == 0) {
            q0[0] = src;
            s_n = 1;
            s_nfar = 0;
            s_thr = delta;
        }
    } else {  // resume: queue parity, near-queue size, far-list size, bucket bound
        par = (int)st[0];
        if (threadIdx.x == 0) {
            s_n = (int)st[1];
            s_nfar = (int)st[2];
            s_thr = st[3];
        }
    }
    int *qa = par ? q1 : q0, *qb = par ? q0 : q1;
    int done = 0;  // levels run in this launch
    __syncthreads();
    for (;;) {
        // relax the near queue until it empties
        for (;; ++done) {
            const int nq = s_n;
            if (nq == 0) break;
            const long long thr = s_thr;
            thin_slice_check(jobs, t_start, done, &s_stop);
            __syncthreads();  // every thread has read s_n and s_thr; s_stop is set
            if (s_stop) {
                if (threadIdx.x == 0) {
                    st[0] = par;
                    st[1] = nq;
                    st[2] = s_nfar;
                    st[3] = thr;
                    a.stats[THIN_STAT * lane_no + 2] = 0;
                }
                return;
            }
            if (threadIdx.x == 0) s_cnt = 0;
            for (int t0 = 0; t0 < nq; t0 += THIN_BLK)
                thin_tile<true>(
                    a, qa, nq, t0, s_pre, s_beg, s_val, s_wave,
                    [&](int v, int) {
                        clear_bit_wg(inq, v);  // v may be queued again once improved
                        return ld_wg(&row[v]);
                    },
                    [&](const int(&u)[THIN_EW], const int(&w)[THIN_EW], const int(&own)[THIN_EW],
                        const bool(&ok)[THIN_EW]) {
                        int du[THIN_EW];
#pragma unroll
                        for (int j = 0; j < THIN_EW; ++j) du[j] = ld_wg(&row[u[j]]);
#pragma unroll
                        for (int j = 0; j < THIN_EW; ++j) {
                            const long long nd = (long long)s_val[own[j]] + w[j];
                            bool near = false, faraway = false;
                            if (ok[j] && nd < du[j]) {
                                const int old = fetch_min_wg(&row[u[j]], (int)nd);
                                if (nd < old) {
                                    if (nd < thr) near = set_bit_first_wg(inq, u[j]);
                                    else faraway = set_bit_first_wg(farbm, u[j]);
                                }
                            }
                            thin_append(near, u[j], qb, &s_cnt);
                            thin_append(faraway, u[j], far, &s_nfar);
                        }
                    });
            __syncthreads();
            if (threadIdx.x == 0) s_n = s_cnt;
            int *t = qa;
            qa = qb;
            qb = t;
            par ^= 1;
            __syncthreads();
        }
        // next bucket: the smallest waiting distance at or above the old bound
        const int nfar = s_nfar;
        const long long old_thr = s_thr;
        long long lo = LLONG_MAX;
        for (int i = threadIdx.x; i < nfar; i += THIN_BLK) {
            const long long d = ld_wg(&row[ld_wg(&far[i])]);
            if (d >= old_thr) lo = d < lo ? d : lo;
        }
        lo = block_min<THIN_BLK>(lo);
        if (lo == LLONG_MAX) break;  // nothing waits: done
        const long long thr = (lo / delta + 1) * delta;
        // stable in-place compaction of the entries still at or above the new bound;
        // the others leave the far list and, unless stale, join the near queue
        if (threadIdx.x == 0) {
            s_cnt = 0;
            s_kept = 0;
        }
        __syncthreads();
        for (int b = 0; b < nfar; b += THIN_BLK) {
            const int i = b + (int)threadIdx.x;
            const int u = i < nfar ? ld_wg(&far[i]) : -1;
            bool keep = false, join = false;
            if (u >= 0) {
                const long long d = ld_wg(&row[u]);
                keep = d >= thr;
                if (!keep) {
                    clear_bit_wg(farbm, u);
                    join = d >= old_thr && set_bit_first_wg(inq, u);
                }
            }
            int total;
            const int ex = block_exclusive_scan<THIN_BLK>(keep ? 1 : 0, s_wave, &total);
            const int kept = s_kept;
            if (keep) far[kept + ex] = u;  // kept + ex <= i: no unread entry is overwritten
            thin_append(join, u, qa, &s_cnt);
            __syncthreads();
            if (threadIdx.x == 0) s_kept = kept + total;
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            s_thr = thr;
            s_nfar = s_kept;
            s_n = s_cnt;
        }
        __syncthreads();
    }
    thin_finish(a, jobs, row, INT_MAX, lane_no);
}

int64_t thin_dev_bytes(const mg_graph *g, bool sssp, int lanes) {
    const int64_t n = g->n, nwords = ((g->n + 63) / 64) * 2;
    const int64_t per = 4 * 3 * n + (sssp ? 4 * n + 8 * nwords : 0) + 8 * THIN_STAT + 32;
    return per * lanes;
}

void thin_dev_free(ThinDev &t) {
    free_dev(t.a.rows);
    free_dev(t.a.q);
    free_dev(t.a.far);
    free_dev(t.a.bits);
    free_dev(t.a.state);
    if (t.h_stats) (void)hipHostFree(t.h_stats);  // a.stats is this pinned array
    t = ThinDev{};
}

int thin_dev_alloc(ThinDev &t, mg_graph *g, bool sssp, int lanes, long long delta) {
    thin_dev_free(t);
    const size_t n = (size_t)g->n;
    t.lanes = lanes;
    t.a.off = g->row_off;
    t.a.adj = g->col;
    t.a.w = g->w;
    t.a.n = g->n;
    t.a.nwords = ((g->n + 63) / 64) * 2;
    t.a.delta = delta;
    int rc;
    if ((rc = alloc_i32(&t.a.rows, lanes * n)) || (rc = alloc_i32(&t.a.q, lanes * 2 * n)) ||
        (rc = alloc_i32((int32_t **)&t.a.state, lanes * 8))) {
        thin_dev_free(t);
        return rc;
    }
    if (sssp && ((rc = alloc_i32(&t.a.far, lanes * n)) ||
                 (rc = alloc_i32((int32_t **)&t.a.bits, lanes * 2 * (size_t)t.a.nwords)))) {
        thin_dev_free(t);
        return rc;
    }
    // each workgroup writes its query's counts straight into pinned host memory; they
    // are read once the launch's event has completed
    if (hipHostMalloc((void **)&t.h_stats, 8 * THIN_STAT * (size_t)lanes, hipHostMallocDefault) != hipSuccess) {
        thin_dev_free(t);
        return set_error(MG_ERR_OOM, "thin search state allocation failed");
    }
    t.a.stats = t.h_stats;
    t.bytes = thin_dev_bytes(g, sssp, lanes);
    return MG_OK;
}

int thin_launch(const ThinArgs &a, const ThinBatch &jobs, bool sssp, hipStream_t st) {
    if (jobs.n <= 0) return MG_OK;
    if (sssp)
        hipLaunchKernelGGL(thin_sssp, dim3(jobs.n), dim3(THIN_BLK), 0, st, a, jobs);
    else
        hipLaunchKernelGGL(thin_bfs, dim3(jobs.n), dim3(THIN_BLK), 0, st, a, jobs);
    MG_HIP(hipGetLastError());
    return MG_OK;
}

}  // namespace mg
