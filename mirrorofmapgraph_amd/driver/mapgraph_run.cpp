// mapgraph_run: command-line driver over the C-ABI, mirroring the reference drivers'
// interface (Algorithms/BFS/bfs.cu:240-520, same flags for SSSP/CC/PageRank):
//
//   mapgraph_run <bfs|sssp|cc|pagerank|bfs_pred> -g graph.mtx [-p k=v,...] [-c file]
//                [-s sources.txt | -s RANDOM] [-o out.txt]
//
// -g   Matrix Market graph (general files are not symmetrized, as in the reference
//      loader, market.cuh:42-166; directed=0 walks in- and out-edges)
// -p   parameters (Config names: src, origin, iter_num, threshold, directed, num_src,
//      device, stats, ...); -c a file of such lines
// -s   a file of starting vertices (one per line, origin-adjusted, at most 1000) or
//      RANDOM (num_src vertices with out-degree > 0, seeded by `seed` instead of
//      time(NULL)); without -s the single source `src`
// -o   one value per line for the last source, as the reference writes it
//
// Reports per source the traversed-edge count (bfs.cu:483-493), the device time and
// GTEPS.  Several sources of BFS or SSSP run as batches of up to 256 (row s of a batch
// equals the single-source result).  run_CPU is not available: correctness checks
// against the CPU oracle live in the test suite.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mapgraph/config.h"
#include "mapgraph_c.h"

namespace {

int fail(const char *what) {
    char msg[1024];
    mg_last_error(msg, sizeof msg);
    std::fprintf(stderr, "%s failed: %s\n", what, msg);
    return 1;
}

void usage(const char *prog) {
    std::printf("usage: %s <bfs|sssp|cc|pagerank|bfs_pred> -g graph.mtx [-p k=v,...] [-c file] "
                "[-s sources.txt|RANDOM] [-o out.txt]\n", prog);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        usage(argv[0]);
        return 1;
    }
    const std::string algo = argv[1];
    if (algo != "bfs" && algo != "sssp" && algo != "cc" && algo != "pagerank" && algo != "bfs_pred") {
        usage(argv[0]);
        return 1;
    }
    const char *graph_file = nullptr, *out_file = nullptr;
    std::string source_file;
    mapgraph::Config cfg;
    cfg.registerParameter<int>("seed", "seed of RANDOM sources", 1);
    try {
        for (int i = 2; i < argc; ++i) {
            const std::string a = argv[i];
            if ((a == "-g" || a == "-graph") && i + 1 < argc) graph_file = argv[++i];
            else if ((a == "-o" || a == "-output") && i + 1 < argc) out_file = argv[++i];
            else if ((a == "-s" || a == "-sources") && i + 1 < argc) source_file = argv[++i];
            else if ((a == "-p" || a == "-parameters") && i + 1 < argc) cfg.parseParameterString(argv[++i]);
            else if (a == "-c" && i + 1 < argc) cfg.parseFile(argv[++i]);
            else if (a == "-help") {
                usage(argv[0]);
                return 0;
            } else {
                usage(argv[0]);
                return 1;
            }
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "bad parameter: %s\n", e.what());
        return 1;
    }
    if (!graph_file) {
        usage(argv[0]);
        return 1;
    }
    const int device = cfg.getParameter<int>("device"), origin = cfg.getParameter<int>("origin");
    if (mg_device_init(device)) return fail("mg_device_init");

    int32_t n = 0, symmetric = 0;
    int64_t nnz = 0;
    if (mg_mtx_read_header(graph_file, &n, &nnz, &symmetric)) return fail("mg_mtx_read_header");
    std::vector<int32_t> src(nnz), dst(nnz), val(nnz);
    int64_t m_read = 0;
    if (mg_mtx_read(graph_file, src.data(), dst.data(), val.data(), nnz, &m_read)) return fail("mg_mtx_read");
    std::vector<int32_t> ro((size_t)n + 1), ci((size_t)m_read), cv((size_t)m_read);
    int64_t m = 0;
    if (mg_csr_from_edges(n, m_read, src.data(), dst.data(), val.data(), MG_CSR_SORT_COLS, ro.data(), ci.data(),
                          cv.data(), m_read, &m))
        return fail("mg_csr_from_edges");
    mg_graph *g = nullptr;
    if (mg_graph_create(ro.data(), ci.data(), cv.data(), n, m, MG_GRAPH_WEIGHTED | MG_GRAPH_BUILD_CSC, &g))
        return fail("mg_graph_create");
    std::printf("Graph %s: %d vertices, %lld edges\n", graph_file, n, (long long)m);
    if (cfg.getParameter<int>("stats")) {
        int64_t dmax = 0, zero = 0;
        for (int v = 0; v < n; ++v) {
            const int64_t d = ro[v + 1] - ro[v];
            dmax = d > dmax ? d : dmax;
            zero += d == 0;
        }
        std::printf("out-degree: max %lld, zero %lld, mean %.2f\n", (long long)dmax, (long long)zero,
                    n ? (double)m / n : 0.0);
    }

    mg_options o;
    mg_options_default(&o);
    o.iter_num = cfg.getParameter<int>("iter_num");
    o.threshold = cfg.getParameter<int>("threshold");
    o.directed = cfg.getParameter<int>("directed");
    o.max_queue_sizing = cfg.getParameter<double>("max_queue_sizing");
    o.verbose = cfg.getParameter<int>("verbose");

    // sources (bfs.cu:340-401)
    std::vector<int32_t> srcs;
    if (source_file == "RANDOM") {
        const int k = cfg.getParameter<int>("num_src");
        srcs.resize(k > 0 ? k : 1);
        if (mg_pick_sources(n, ro.data(), (int32_t)srcs.size(), (uint64_t)cfg.getParameter<int>("seed"), srcs.data()))
            return fail("mg_pick_sources");
        std::printf("Using %zu random starting vertices!\n", srcs.size());
    } else if (!source_file.empty()) {
        FILE *f = std::fopen(source_file.c_str(), "r");
        if (!f) {
            std::fprintf(stderr, "Source file open error!\n");
            return 1;
        }
        int v;
        while (srcs.size() < 1000 && std::fscanf(f, "%d", &v) == 1) srcs.push_back(origin == 1 ? v - 1 : v);
        std::fclose(f);
        std::printf("number of srcs used: %zu\n", srcs.size());
    } else {
        const int s0 = cfg.getParameter<int>("src");
        srcs.push_back(origin == 1 ? s0 - 1 : s0);
        std::printf("Single source vertex: %d\n", srcs[0]);
    }

    std::vector<int32_t> out((size_t)n), pred;
    std::vector<float> rank;
    mg_stats st;
    double total_ms = 0;
    long long total_te = 0;
    int rc = 0;
    auto report = [&](int s, const mg_stats &x) {
        std::printf("src %d: iterations %d, traversed edges %lld, kernel time %.3f ms, %.3f GTEPS\n", s,
                    x.iterations, (long long)x.edges_traversed, x.device_ms,
                    x.device_ms > 0 ? x.edges_traversed / x.device_ms / 1e6 : 0.0);
    };
    if (algo == "cc" || algo == "pagerank") {
        if (algo == "cc") rc = mg_cc(g, out.data(), &o, &st);
        else {
            rank.resize((size_t)n);
            rc = mg_pagerank(g, rank.data(), &o, &st);
        }
        if (rc) return fail(algo.c_str());
        total_ms = st.device_ms;
        total_te = st.edges_traversed;
        std::printf("iterations %d, kernel time %.3f ms\n", st.iterations, st.device_ms);
    } else if (algo == "bfs_pred") {
        pred.resize((size_t)n);
        for (int32_t s : srcs) {
            if ((rc = mg_bfs_pred(g, s, out.data(), pred.data(), &o, &st))) return fail("mg_bfs_pred");
            report(s, st);
            total_ms += st.device_ms;
            total_te += st.edges_traversed;
        }
    } else if (srcs.size() == 1) {
        rc = algo == "bfs" ? mg_bfs(g, srcs[0], out.data(), &o, &st) : mg_sssp(g, srcs[0], out.data(), &o, &st);
        if (rc) return fail(algo.c_str());
        report(srcs[0], st);
        total_ms = st.device_ms;
        total_te = st.edges_traversed;
    } else {
        // several sources: batches of up to MG_BATCH_MAX, one device-driven search each
        std::vector<int32_t> rows;
        for (size_t b = 0; b < srcs.size(); b += MG_BATCH_MAX) {
            const int k = (int)std::min<size_t>(MG_BATCH_MAX, srcs.size() - b);
            rows.resize((size_t)k * n);
            rc = algo == "bfs" ? mg_bfs_batch(g, srcs.data() + b, k, rows.data(), &o, &st)
                               : mg_sssp_batch(g, srcs.data() + b, k, rows.data(), &o, &st);
            if (rc) return fail(algo.c_str());
            std::printf("batch of %d sources: traversed edges %lld, kernel time %.3f ms, %.3f GTEPS\n", k,
                        (long long)st.edges_traversed, st.device_ms, st.edges_traversed / st.device_ms / 1e6);
            total_ms += st.device_ms;
            total_te += st.edges_traversed;
            std::memcpy(out.data(), rows.data() + (size_t)(k - 1) * n, 4 * (size_t)n);
        }
    }
    std::printf("Traversed edge: %lld\n", total_te);
    std::printf("Kernel time took: %.3f ms, %.3f GTEPS\n", total_ms, total_ms > 0 ? total_te / total_ms / 1e6 : 0.0);
    if (out_file) {
        FILE *f = std::fopen(out_file, "w");
        if (!f) {
            std::fprintf(stderr, "cannot open %s\n", out_file);
            return 1;
        }
        for (int v = 0; v < n; ++v) {
            if (algo == "pagerank") std::fprintf(f, "%.9g\n", rank[v]);
            else if (algo == "bfs_pred") std::fprintf(f, "%d %d\n", out[v], pred[v]);
            else std::fprintf(f, "%d\n", out[v]);
        }
        std::fclose(f);
    }
    mg_graph_destroy(g);
    return 0;
}
