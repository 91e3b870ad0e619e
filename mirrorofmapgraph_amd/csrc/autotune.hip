// Autotuner: picks the runtime parameters of one search kind for one graph.
//
// Not in the reference, which fixes its kernel policies per SM version at compile
// time (GASengine/enactor_vertex_centric.cuh:2910-2966) and takes one `threshold`
// for its expand strategy (Algorithms/BFS/register.h:39).  Here the parameters that
// decide how a search maps onto the GPU are mg_options fields, so they can be chosen
// per graph: the direction switches of BFS (alpha, beta) and of the batched BFS
// (batch_alpha, batch_beta), the SSSP bucket width and bucket lift point
// (sssp_delta, sssp_bucket_edges), and the workgroups of the whole-GPU level kernels
// (level_grid).  The tuner runs coordinate descent over candidate values: each
// candidate setting is timed on the caller's sample sources (device time, median of
// `reps` runs after a warm-up run) and kept when it is at least 3% faster than the
// best so far; two passes over the parameters.  Every candidate's results must be
// bit-identical to the starting setting's (they are unique answers), otherwise the
// tuner stops with MG_ERR_DEVICE: a parameter must never change a result.
#include <algorithm>
#include <cstring>
#include <vector>

#include "engine.h"

namespace mg {
namespace {

struct Param {
    int32_t mg_options::*field;
    std::vector<int32_t> values;
};

struct Tuner {
    mg_graph *g;
    int what;
    const int32_t *srcs;
    int nsrc;
    int reps;
    std::vector<std::vector<int32_t>> ref;  // results of the starting setting, per run unit
    int candidates = 0;

    size_t row_ints() const {
        return (size_t)g->n * ((what == MG_TUNE_BFS_BATCH || what == MG_TUNE_SSSP_BATCH) ? nsrc : 1);
    }
    int units() const { return (what == MG_TUNE_BFS || what == MG_TUNE_SSSP) ? nsrc : 1; }

    // One search of unit u (a source, or the whole batch) with output to `out`
    // (nullptr: left on the device); device ms in *ms.
    int once(const mg_options &o, int u, int32_t *out, double *ms) {
        mg_stats st;
        int rc;
        switch (what) {
            case MG_TUNE_BFS: rc = mg_bfs(g, srcs[u], out, &o, &st); break;
            case MG_TUNE_SSSP: rc = mg_sssp(g, srcs[u], out, &o, &st); break;
            case MG_TUNE_BFS_BATCH: rc = mg_bfs_batch(g, srcs, nsrc, out, &o, &st); break;
            default: rc = mg_sssp_batch(g, srcs, nsrc, out, &o, &st); break;
        }
        *ms = st.device_ms;
        return rc;
    }

    // Times setting o (sum over units of the median of reps runs) and checks its
    // results against the starting setting's (recorded when ref is empty).
    int measure(const mg_options &o, double *total) {
        ++candidates;
        *total = 0;
        const bool record = ref.empty();
        std::vector<int32_t> buf(row_ints());
        for (int u = 0; u < units(); ++u) {
            double ms;
            int rc = once(o, u, buf.data(), &ms);  // warm-up (captures step graphs) and the check
            if (rc) return rc;
            if (record) {
                ref.push_back(buf);
            } else if (std::memcmp(buf.data(), ref[u].data(), 4 * buf.size()) != 0) {
                return set_error(MG_ERR_DEVICE, "mg_autotune: a candidate setting changed the results (unit %d)", u);
            }
            std::vector<double> t;
            for (int r = 0; r < reps; ++r) {
                if ((rc = once(o, u, nullptr, &ms))) return rc;
                t.push_back(ms);
            }
            std::sort(t.begin(), t.end());
            *total += t[t.size() / 2];
        }
        return MG_OK;
    }
};

std::vector<Param> params_for(const mg_graph *g, int what) {
    const int32_t m8 = (int32_t)std::min<int64_t>(INT32_MAX, std::max<int64_t>(1, g->m / 8));
    auto frac = [&](int64_t d) { return (int32_t)std::min<int64_t>(INT32_MAX, std::max<int64_t>(1, g->m / d)); };
    const int32_t w4 = std::max(1, 4 * g->max_weight);
    switch (what) {
        case MG_TUNE_BFS:
            return {{&mg_options::alpha, {4, 8, 14, 24, 48}},
                    {&mg_options::beta, {6, 12, 24, 48}},
                    {&mg_options::level_grid, {0, 512, 1024, 4096, 8192}}};
        case MG_TUNE_SSSP:
            return {{&mg_options::sssp_delta, {-1, 1, 2, 4, 8, 16, 64, w4}},
                    {&mg_options::sssp_bucket_edges, {m8, frac(32), frac(4), frac(2), -1}},
                    {&mg_options::level_grid, {0, 2048, 4096, 16384}}};
        case MG_TUNE_BFS_BATCH:
            return {{&mg_options::batch_alpha, {2, 4, 8, 16}}, {&mg_options::batch_beta, {4, 8, 16, 32, 64}}};
        default:
            return {{&mg_options::sssp_delta, {1, 2, 4, 8, 16, 32, 64, 128, w4}}};
    }
}

}  // namespace
}  // namespace mg

using namespace mg;

extern "C" int mg_autotune(mg_graph *g, int32_t what, const int32_t *srcs, int32_t nsrc, int32_t reps,
                           const mg_options *start, mg_options *best, mg_tune_report *report) {
    clear_error();
    if (!g || !best) return set_error(MG_ERR_INVALID_ARG, "mg_autotune: null graph or result");
    if (what < MG_TUNE_BFS || what > MG_TUNE_SSSP_BATCH)
        return set_error(MG_ERR_INVALID_ARG, "mg_autotune: unknown search kind %d", what);
    const int maxsrc = (what == MG_TUNE_BFS || what == MG_TUNE_SSSP) ? 64 : MG_BATCH_MAX;
    if (!srcs || nsrc < 1 || nsrc > maxsrc)
        return set_error(MG_ERR_INVALID_ARG, "mg_autotune: 1 to %d sample sources needed", maxsrc);
    if (reps < 1 || reps > 50) return set_error(MG_ERR_INVALID_ARG, "mg_autotune: reps must be in [1, 50]");
    for (int i = 0; i < nsrc; ++i)
        if (srcs[i] < 0 || srcs[i] >= g->n)
            return set_error(MG_ERR_INVALID_ARG, "mg_autotune: source %d outside [0,%d)", srcs[i], g->n);
    mg_options cur;
    if (start) cur = *start;
    else mg_options_default(&cur);
    cur.out_on_device = 0;
    cur.verbose = 0;
    Tuner t{g, what, srcs, nsrc, reps, {}};
    double t_start = 0, t_best = 0;
    int rc = t.measure(cur, &t_start);
    if (rc) return rc;
    t_best = t_start;
    const std::vector<Param> ps = params_for(g, what);
    for (int pass = 0; pass < 2; ++pass) {
        bool moved = false;
        for (const Param &p : ps) {
            for (int32_t v : p.values) {
                if (cur.*p.field == v) continue;
                mg_options o = cur;
                o.*p.field = v;
                double ms = 0;
                if ((rc = t.measure(o, &ms))) return rc;
                if (ms < 0.97 * t_best) {
                    t_best = ms;
                    cur = o;
                    moved = true;
                }
            }
        }
        if (!moved) break;
    }
    *best = cur;
    if (report) {
        std::memset(report, 0, sizeof *report);
        report->start_ms = t_start;
        report->best_ms = t_best;
        report->candidates = t.candidates;
    }
    return MG_OK;
}
