// host_tools_check -- drives the host half of the library (graph_host.cpp: Matrix
// Market reader, generators, CSR / CSC builders, weights, source picker, their
// argument checks) and the CPU oracle (mg_oracle.c) through valid and invalid inputs,
// for an AddressSanitizer + UndefinedBehaviorSanitizer build on the host.  Every
// check that fails prints a line; the exit status is the number of failures.
//
// usage: host_tools_check <fixture dir> <scratch dir>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mapgraph_c.h"

extern "C" {
int oracle_bfs(int directed, int n, const int *row_offsets, const int *col_idx, const int *col_offsets,
               const int *row_idx, int src, int *depth);
void oracle_sssp_quadratic(int directed, int n, const int *row_offsets, const int *col_idx, const int *weights,
                           const int *col_offsets, const int *row_idx, const int *csc_edge_id, int src, int unreached,
                           int *dist);
void oracle_sssp_heap(int directed, int n, const int *row_offsets, const int *col_idx, const int *weights,
                      const int *col_offsets, const int *row_idx, const int *csc_edge_id, int src, int unreached,
                      int *dist);
int oracle_cc(int n, const int *row_offsets, const int *col_idx, const int *col_offsets, const int *row_idx,
              int *label);
void oracle_cc_union_find(int n, const int *row_offsets, const int *col_idx, int *label);
int oracle_pr_gauss_seidel(int n, const int *row_offsets, const int *col_offsets, const int *row_idx, int max_sweeps,
                           float *rank);
int oracle_pr_gas(int n, const int *row_offsets, const int *col_idx, const int *col_offsets, const int *row_idx,
                  int iter_num, float teleport, float eps, float *rank);
int oracle_pr_jacobi_f32(int n, const int *row_offsets, const int *col_offsets, const int *row_idx, int iters,
                         float teleport, double tol, float *rank, double *l1);
int oracle_pr_jacobi_f64(int n, const int *row_offsets, const int *col_offsets, const int *row_idx, int iters,
                         double teleport, double tol, double *rank, double *l1);
long long oracle_teps_edges(int n, const int *row_offsets, const int *label, int unreached);
}

static int failures = 0;
#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);        \
            ++failures;                                                          \
        }                                                                        \
    } while (0)

struct Csr {
    int n = 0;
    std::vector<int32_t> ro, col, val;
    std::vector<int32_t> co, ri, eid;  // CSC
    int64_t m() const { return (int64_t)col.size(); }
};

static Csr build(int n, const std::vector<int32_t> &s, const std::vector<int32_t> &d, const std::vector<int32_t> *v,
                 uint32_t flags) {
    Csr g;
    g.n = n;
    const int64_t cap = (flags & MG_CSR_SYMMETRIZE) ? 2 * (int64_t)s.size() : (int64_t)s.size();
    g.ro.assign(n + 1, 0);
    g.col.assign(cap > 0 ? cap : 1, 0);
    if (v) g.val.assign(cap > 0 ? cap : 1, 0);
    int64_t m = 0;
    int rc = mg_csr_from_edges(n, (int64_t)s.size(), s.data(), d.data(), v ? v->data() : nullptr, flags, g.ro.data(),
                               g.col.data(), v ? g.val.data() : nullptr, cap, &m);
    CHECK(rc == MG_OK);
    g.col.resize(m);
    if (v) g.val.resize(m);
    g.co.assign(n + 1, 0);
    g.ri.assign(m > 0 ? m : 1, 0);
    g.eid.assign(m > 0 ? m : 1, 0);
    CHECK(mg_csc_from_csr(n, m, g.ro.data(), g.col.data(), g.co.data(), g.ri.data(), g.eid.data()) == MG_OK);
    // the degree-order renaming: a permutation, rows of the renamed graph non-increasing in degree
    std::vector<int32_t> perm(n > 0 ? n : 1), ro2(n + 1), col2(m > 0 ? m : 1), val2(m > 0 ? m : 1);
    CHECK(mg_csr_relabel_degree(n, m, g.ro.data(), g.col.data(), v ? g.val.data() : nullptr, perm.data(), ro2.data(),
                                col2.data(), v ? val2.data() : nullptr) == MG_OK);
    std::vector<char> seen(n > 0 ? n : 1, 0);
    for (int i = 0; i < n; ++i) {
        CHECK(perm[i] >= 0 && perm[i] < n && !seen[perm[i]]);
        if (perm[i] >= 0 && perm[i] < n) seen[perm[i]] = 1;
    }
    for (int i = 0; i + 1 < n; ++i) CHECK(ro2[i + 1] - ro2[i] >= ro2[i + 2] - ro2[i + 1]);
    return g;
}

static void run_oracle(const Csr &g) {
    const int n = g.n;
    std::vector<int> a(n > 0 ? n : 1), b(n > 0 ? n : 1);
    std::vector<int32_t> w(g.m() > 0 ? g.m() : 1, 1);
    if (!g.val.empty()) w = g.val;
    for (int src = 0; src < n && src < 3; ++src)
        for (int directed = 0; directed < 2; ++directed) {
            oracle_bfs(directed, n, g.ro.data(), g.col.data(), g.co.data(), g.ri.data(), src, a.data());
            oracle_sssp_heap(directed, n, g.ro.data(), g.col.data(), w.data(), g.co.data(), g.ri.data(),
                             g.eid.data(), src, INT_MAX, a.data());
            oracle_sssp_quadratic(directed, n, g.ro.data(), g.col.data(), w.data(), g.co.data(), g.ri.data(),
                                  g.eid.data(), src, INT_MAX, b.data());
            CHECK(a == b);
            (void)oracle_teps_edges(n, g.ro.data(), a.data(), INT_MAX);
        }
    oracle_cc(n, g.ro.data(), g.col.data(), g.co.data(), g.ri.data(), a.data());
    oracle_cc_union_find(n, g.ro.data(), g.col.data(), b.data());
    CHECK(a == b);
    std::vector<float> r(n > 0 ? n : 1);
    std::vector<double> r64(n > 0 ? n : 1);
    double l1 = 0;
    (void)oracle_pr_gauss_seidel(n, g.ro.data(), g.co.data(), g.ri.data(), 100, r.data());
    (void)oracle_pr_gas(n, g.ro.data(), g.col.data(), g.co.data(), g.ri.data(), 100, 0.15f, 0.01f, r.data());
    (void)oracle_pr_jacobi_f32(n, g.ro.data(), g.co.data(), g.ri.data(), 20, 0.15f, 1e-6, r.data(), &l1);
    (void)oracle_pr_jacobi_f64(n, g.ro.data(), g.co.data(), g.ri.data(), 20, 0.15, 1e-6, r64.data(), &l1);
}

static void write_file(const std::string &path, const char *text) {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) {
        std::printf("FAILED cannot write %s\n", path.c_str());
        ++failures;
        return;
    }
    std::fputs(text, f);
    std::fclose(f);
}

static int read_mtx(const std::string &path, Csr *out) {
    int32_t n = 0, sym = 0;
    int64_t nnz = 0;
    int rc = mg_mtx_read_header(path.c_str(), &n, &nnz, &sym);
    if (rc) return rc;
    std::vector<int32_t> s(nnz > 0 ? nnz : 1), d(nnz > 0 ? nnz : 1), v(nnz > 0 ? nnz : 1);
    int64_t got = 0;
    rc = mg_mtx_read(path.c_str(), s.data(), d.data(), v.data(), nnz, &got);
    if (rc) return rc;
    s.resize(got);
    d.resize(got);
    v.resize(got);
    if (out) *out = build(n, s, d, &v, 0);
    return MG_OK;
}

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <fixture dir> <scratch dir>\n", argv[0]);
        return 100;
    }
    const std::string fix = argv[1], tmp = argv[2];
    char msg[256];

    // fixtures through the loader and every oracle routine
    for (const char *name : {"small.mtx", "trivialGraph.mtx", "chesapeake.mtx"}) {
        Csr g;
        CHECK(read_mtx(fix + "/" + name, &g) == MG_OK);
        run_oracle(g);
    }
    // malformed and missing Matrix Market files: status codes, no crash
    struct Bad {
        const char *text;
        int code;
    } bad[] = {
        {"%%MatrixMarket matrix coordinate pattern general\n3 4 2\n1 2\n", MG_ERR_FORMAT},  // too few entries
        {"3 3 2\n1 2\n", MG_ERR_FORMAT},
        {"3 3 1\n1 4\n", MG_ERR_FORMAT},                                                    // id outside [1, n]
        {"3 3 1\n0 1\n", MG_ERR_FORMAT},                                                    // 0 in a 1-based file
        {"3 3 1\nx y\n", MG_ERR_FORMAT},                                                    // not numbers
        {"", MG_ERR_FORMAT},                                                                // empty
        {"%%MatrixMarket matrix coordinate real general\n-3 3 1\n1 1 2.5\n", MG_ERR_FORMAT},  // negative size
        {"2 2 3\n1 2\n2 1\n1 1\n2 2\n", MG_ERR_FORMAT},                                      // more than declared
    };
    for (size_t i = 0; i < sizeof bad / sizeof bad[0]; ++i) {
        const std::string p = tmp + "/bad" + std::to_string(i) + ".mtx";
        write_file(p, bad[i].text);
        const int rc = read_mtx(p, nullptr);
        if (rc == MG_OK) {
            std::printf("FAILED malformed file %zu was accepted\n", i);
            ++failures;
        }
        (void)mg_last_error(msg, sizeof msg);
    }
    CHECK(read_mtx(tmp + "/missing.mtx", nullptr) == MG_ERR_IO);

    // generators and builders, including the empty and tiny cases
    for (int scale = 1; scale <= 10; scale += 3) {
        const int64_t m = mg_rmat_num_edges(scale, 8);
        std::vector<int32_t> s(m), d(m);
        CHECK(mg_gen_rmat_edges(scale, 8, 0.57, 0.19, 0.19, 42, s.data(), d.data()) == MG_OK);
        for (uint32_t flags = 0; flags < 16; ++flags) {
            std::vector<int32_t> v(m);
            for (int64_t k = 0; k < m; ++k) v[k] = (int32_t)(k % 13);
            Csr g = build(1 << scale, s, d, &v, flags);
            std::vector<int32_t> w(g.m() > 0 ? g.m() : 1);
            CHECK(mg_edge_weights_hash(g.n, g.ro.data(), g.col.data(), g.m(), 7, 64, w.data()) == MG_OK);
            if (flags == 15) run_oracle(g);
        }
        std::vector<int32_t> srcs(17);
        Csr g = build(1 << scale, s, d, nullptr, MG_CSR_SYMMETRIZE | MG_CSR_DEDUPE | MG_CSR_DROP_LOOPS);
        (void)mg_pick_sources(g.n, g.ro.data(), 17, 3, srcs.data());
    }
    for (int wdt = 1; wdt <= 9; wdt += 4) {
        const int64_t E = mg_grid_num_edges(wdt, wdt + 1);
        std::vector<int32_t> s(E > 0 ? E : 1), d(E > 0 ? E : 1);
        int64_t m = 0;
        CHECK(mg_gen_grid_edges(wdt, wdt + 1, 0.05, 9, s.data(), d.data(), &m) == MG_OK);
        s.resize(m);
        d.resize(m);
        run_oracle(build(wdt * (wdt + 1), s, d, nullptr, MG_CSR_SYMMETRIZE | MG_CSR_SORT_COLS));
    }
    Csr empty = build(0, {}, {}, nullptr, MG_CSR_SYMMETRIZE);
    run_oracle(empty);
    run_oracle(build(5, {}, {}, nullptr, 0));

    // argument checks: bad sizes, ids and null pointers give status codes
    int32_t ro[3] = {0, 1, 2}, col[2] = {1, 0}, out[4];
    int64_t m_out = 0;
    int32_t src_bad[1] = {7}, dst_ok[1] = {0};
    CHECK(mg_csr_from_edges(2, 1, src_bad, dst_ok, nullptr, 0, ro, col, nullptr, 2, &m_out) != MG_OK);
    CHECK(mg_csr_from_edges(-1, 0, nullptr, nullptr, nullptr, 0, ro, col, nullptr, 0, &m_out) != MG_OK);
    CHECK(mg_csr_from_edges(2, 1, nullptr, dst_ok, nullptr, 0, ro, col, nullptr, 2, &m_out) != MG_OK);
    int32_t s2[2] = {0, 1}, d2[2] = {1, 0};
    CHECK(mg_csr_from_edges(2, 2, s2, d2, nullptr, MG_CSR_SYMMETRIZE, ro, col, nullptr, 2, &m_out) != MG_OK);  // capacity
    CHECK(mg_gen_rmat_edges(-1, 8, 0.57, 0.19, 0.19, 1, out, out) != MG_OK);
    CHECK(mg_gen_rmat_edges(3, 8, 0.9, 0.19, 0.19, 1, nullptr, nullptr) != MG_OK);
    CHECK(mg_gen_grid_edges(-2, 3, 0.05, 1, out, out, &m_out) != MG_OK);
    CHECK(mg_edge_weights_hash(2, ro, col, 2, 1, 0, out) != MG_OK);
    CHECK(mg_pick_sources(2, ro, -1, 1, out) != MG_OK);
    CHECK(mg_csc_from_csr(2, 2, nullptr, col, out, out, out) != MG_OK);
    CHECK(mg_csr_relabel_degree(2, 2, nullptr, col, nullptr, out, out, out, nullptr) != MG_OK);
    CHECK(mg_mtx_read_header(nullptr, nullptr, nullptr, nullptr) != MG_OK);
    CHECK(mg_last_error(msg, sizeof msg) > 0);
    CHECK(mg_last_error(nullptr, 0) > 0);

    std::printf("host_tools_check: %d failure(s)\n", failures);
    return failures;
}
