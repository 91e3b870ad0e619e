// The generic GAS enactor (include/mapgraph/enactor.h) instantiated for the four
// ported MapGraph programs and reachable through the C-ABI with
// mg_options.engine = MG_ENGINE_GAS.  The program state (CsrProblem wrapping the
// graph's device arrays, enactor buffers, the program's vertex arrays) is kept per
// graph and reused by later searches.
#include <memory>
#include <string>

#include "../../include/mapgraph/enactor.h"
#include "../../include/mapgraph/programs/bfs.h"
#include "../../include/mapgraph/programs/cc.h"
#include "../../include/mapgraph/programs/pagerank.h"
#include "../../include/mapgraph/programs/sssp.h"
#include "engine.h"

namespace mg {

struct GasStateBase {
    virtual ~GasStateBase() = default;
};

template <class Program>
struct GasState : GasStateBase {
    typedef typename Program::DataType EValue;
    typedef GASengine::CsrProblem<Program, int, int, EValue, false, true> Problem;
    typedef GASengine::EnactorVertexCentric<Problem, Program, false> Enactor;
    mapgraph::Config cfg;
    Problem problem;
    Enactor enactor;
    explicit GasState(const mapgraph::Config &c) : cfg(c), problem(c), enactor(c) {}
};

template <class Program>
static int gas_run(mg_graph *g, int kind, const int *srcs, int nsrc, typename Program::DataType *out,
                   const mg_options &o, mg_stats *stats, const CallTimer &timer) {
    typedef GasState<Program> S;
    if (o.out_on_device)
        return set_error(MG_ERR_UNSUPPORTED, "MG_ENGINE_GAS returns results to host memory only");
    auto &slot = g->steps->gas[kind];
    if (!slot) {
        mapgraph::Config cfg;
        cfg.setParameter("device", std::to_string(g->device));
        auto *s = new (std::nothrow) S(cfg);
        if (!s) return set_error(MG_ERR_OOM, "host allocation failed");
        hipError_t e = s->problem.FromDeviceGraph(g->device, g->n, (int)g->m, g->col, g->row_off,
                                                  (typename S::EValue *)(void *)g->w, g->col_off, g->row_idx,
                                                  g->csc_eid, (g->flags & MG_GRAPH_SYMMETRIC) != 0, g->stream);
        if (e != hipSuccess) {
            delete s;
            return set_error(MG_ERR_DEVICE, "CsrProblem::FromDeviceGraph failed: %s", hipGetErrorString(e));
        }
        slot = s;
    }
    S *s = static_cast<S *>(slot);
    hipError_t e = s->enactor.EnactIterativeSearch(s->problem, nullptr, o.directed, nsrc, const_cast<int *>(srcs),
                                                   o.iter_num, o.threshold);
    if (e == hipErrorInvalidValue)
        return set_error(MG_ERR_UNSUPPORTED, "the program needs the in-edge view (MG_GRAPH_BUILD_CSC or MG_GRAPH_SYMMETRIC)");
    if (e != hipSuccess) return set_error(MG_ERR_DEVICE, "GAS enactor failed: %s", hipGetErrorString(e));
    if (out) {
        e = s->problem.ExtractResults(out);
        if (e != hipSuccess) return set_error(MG_ERR_DEVICE, "ExtractResults failed: %s", hipGetErrorString(e));
    }
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->iterations = s->enactor.iterations();
        stats->device_ms = s->enactor.kernel_ms();
        stats->host_ms = timer.ms();
    }
    return MG_OK;
}

int gas_bfs(mg_graph *g, int src, int *out, const mg_options &o, mg_stats *st, const CallTimer &t) {
    return gas_run<bfs>(g, 0, &src, 1, out, o, st, t);
}
int gas_sssp(mg_graph *g, int src, int *out, const mg_options &o, mg_stats *st, const CallTimer &t) {
    return gas_run<sssp>(g, 1, &src, 1, out, o, st, t);
}
int gas_cc(mg_graph *g, int *out, const mg_options &o, mg_stats *st, const CallTimer &t) {
    return gas_run<cc>(g, 2, nullptr, 0, out, o, st, t);
}
int gas_pagerank(mg_graph *g, float *out, const mg_options &o, mg_stats *st, const CallTimer &t) {
    return gas_run<pagerank>(g, 3, nullptr, 0, out, o, st, t);
}

void gas_release(mg_graph *g) {
    for (auto &p : g->steps->gas) {
        delete p;
        p = nullptr;
    }
}

}  // namespace mg
