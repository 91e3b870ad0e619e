// mapgraph_dropin -- a MapGraph driver written against the C++ drop-in headers
// (include/mapgraph/*.h) with the reference drivers' call sequence
// (Algorithms/BFS/bfs.cu:428-455, Algorithms/SSSP/sssp.cu:378-428, Algorithms/CC/
// cc.cu:300-347, Algorithms/PageRank/PageRank.cu:296-313):
//
//   Config -> CsrProblem<Program,...>(cfg) -> FromHostProblem(...)
//          -> EnactorVertexCentric<CsrProblem, Program, INSTRUMENT>(cfg, DEBUG)
//          -> EnactIterativeSearch(csr_problem, h_row_offsets, directed, num_srcs, srcs,
//                                  iter_num, threshold)
//          -> CsrProblem::ExtractResults(h_values)
//
// Test harness: the GPU tests build a graph on the host, write it as a binary CSR,
// run this program and compare the values it writes with the CPU oracle.
//
// usage: mapgraph_dropin <bfs|sssp|cc|pagerank> <graph.bin> <out.bin> [-p k=v ...]
//   graph.bin: int64 n, int64 m, int64 has_values, int32 row_offsets[n+1],
//              int32 column_indices[m], int32 edge_values[m] when has_values
//   out.bin:   the program's DataType per vertex (int32 labels / distances, float ranks)
// Exit status 0 on success, 1 on bad arguments or input, 2 when a call fails.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <mapgraph/enactor.h>
#include <mapgraph/programs/bfs.h>
#include <mapgraph/programs/cc.h>
#include <mapgraph/programs/pagerank.h>
#include <mapgraph/programs/sssp.h>

namespace {

struct HostGraph {
    int64_t n = 0, m = 0, has_values = 0;
    std::vector<int> row_offsets, column_indices, edge_values;
};

bool read_graph(const char *path, HostGraph &g) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    int64_t hdr[3];
    bool ok = std::fread(hdr, sizeof hdr, 1, f) == 1 && hdr[0] >= 0 && hdr[1] >= 0 && hdr[0] < INT32_MAX &&
              hdr[1] < INT32_MAX;
    if (ok) {
        g.n = hdr[0];
        g.m = hdr[1];
        g.has_values = hdr[2];
        g.row_offsets.resize(g.n + 1);
        g.column_indices.resize(g.m);
        ok = std::fread(g.row_offsets.data(), 4, g.n + 1, f) == (size_t)(g.n + 1) &&
             std::fread(g.column_indices.data(), 4, g.m, f) == (size_t)g.m;
        if (ok && g.has_values) {
            g.edge_values.resize(g.m);
            ok = std::fread(g.edge_values.data(), 4, g.m, f) == (size_t)g.m;
        }
    }
    std::fclose(f);
    return ok;
}

// One program run with the reference driver's sequence; the values go to `out`.
template <class Program>
int run(const mapgraph::Config &cfg, HostGraph &g, const char *out_path) {
    typedef typename Program::DataType Value;
    typedef GASengine::CsrProblem<Program, int, int, Value, false, true> CsrProblem;
    const int directed = cfg.getParameter<int>("directed");
    const int nodes = (int)g.n, edges = (int)g.m;
    std::vector<Value> h_values(g.has_values ? g.m : 0);
    for (int64_t k = 0; k < (int64_t)h_values.size(); ++k) h_values[k] = (Value)g.edge_values[k];

    CsrProblem csr_problem(cfg);
    hipError_t e = csr_problem.FromHostProblem(false, nodes, edges, g.column_indices.data(), g.row_offsets.data(),
                                               h_values.empty() ? nullptr : h_values.data(), nullptr, nullptr,
                                               /*num_gpus*/ 1, directed);
    if (e != hipSuccess) {
        std::fprintf(stderr, "FromHostProblem failed: %s\n", hipGetErrorString(e));
        return 2;
    }
    GASengine::EnactorVertexCentric<CsrProblem, Program, true> vertex_centric(cfg, false);
    // origin=1 (the reference default) makes src 1-based (bfs.cu:392-398)
    const int origin = cfg.getParameter<int>("origin");
    int src[1] = {cfg.getParameter<int>("src") - origin};
    const bool single = Program::srcVertex() == SINGLE;
    if (single && (src[0] < 0 || src[0] >= nodes)) {
        std::fprintf(stderr, "source %d outside [0,%d)\n", src[0], nodes);
        return 1;
    }
    e = vertex_centric.EnactIterativeSearch(csr_problem, g.row_offsets.data(), directed, single ? 1 : 0,
                                            single ? src : nullptr, cfg.getParameter<int>("iter_num"),
                                            cfg.getParameter<int>("threshold"));
    if (e != hipSuccess) {
        std::fprintf(stderr, "EnactIterativeSearch failed: %s\n", hipGetErrorString(e));
        return 2;
    }
    std::vector<Value> h_out(g.n);
    e = csr_problem.ExtractResults(h_out.data());
    if (e != hipSuccess) {
        std::fprintf(stderr, "ExtractResults failed: %s\n", hipGetErrorString(e));
        return 2;
    }
    std::printf("Kernel time took: %f ms\n", vertex_centric.kernel_ms());
    std::printf("Total iteration: %d\n", vertex_centric.iterations());
    FILE *f = std::fopen(out_path, "wb");
    if (!f || std::fwrite(h_out.data(), sizeof(Value), h_out.size(), f) != h_out.size()) {
        if (f) std::fclose(f);
        std::fprintf(stderr, "cannot write %s\n", out_path);
        return 1;
    }
    std::fclose(f);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <bfs|sssp|cc|pagerank> <graph.bin> <out.bin> [-p k=v ...]\n", argv[0]);
        return 1;
    }
    mapgraph::Config cfg;
    for (int i = 4; i + 1 < argc; i += 2) {
        if (std::strcmp(argv[i], "-p") != 0) {
            std::fprintf(stderr, "unknown argument %s\n", argv[i]);
            return 1;
        }
        cfg.parseParameterString(argv[i + 1]);
    }
    HostGraph g;
    if (!read_graph(argv[2], g)) {
        std::fprintf(stderr, "cannot read graph %s\n", argv[2]);
        return 1;
    }
    const std::string prog = argv[1];
    if (prog == "bfs") return run<bfs>(cfg, g, argv[3]);
    if (prog == "sssp") return run<sssp>(cfg, g, argv[3]);
    if (prog == "cc") return run<cc>(cfg, g, argv[3]);
    if (prog == "pagerank") return run<pagerank>(cfg, g, argv[3]);
    std::fprintf(stderr, "unknown program %s\n", argv[1]);
    return 1;
}
