/*
 * mapgraph_c.h -- C-ABI of libmapgraph_hip.so, the MI355X (gfx950) GAS graph engine.
 *
 * Plain pointers and sizes only; no C++ or HIP types cross this boundary, so the
 * library can be bound from C, C++, Python (ctypes) or any FFI.  Every entry point
 * returns an int status (MG_OK == 0); the text of the last error on the calling
 * thread is available from mg_last_error().  No entry point calls exit() or lets an
 * exception escape.
 *
 * Reference interfaces replaced (paths relative to the MapGraph 0.3.3 tree):
 *   mg_graph_create      CsrProblem::FromHostProblem     GASengine/csr_problem.cuh:401-621
 *                        + device CSC build              util.cuh:55-126
 *   mg_graph_destroy     GraphSlice::~GraphSlice         GASengine/csr_problem.cuh:217-289
 *   mg_bfs               EnactIterativeSearch<bfs>       GASengine/enactor_vertex_centric.cuh:2910-2966
 *                        with Algorithms/BFS/bfs.h, then CsrProblem::ExtractResults
 *                        (csr_problem.cuh:371-396) and bfs::extractResult (bfs.h:257-260)
 *   mg_sssp              EnactIterativeSearch<sssp>      Algorithms/SSSP/sssp.h, sssp.cu:378-428
 *   mg_cc                EnactIterativeSearch<cc>        Algorithms/CC/cc.h, cc.cu:300-347
 *   mg_pagerank          EnactIterativeSearch<pagerank>  Algorithms/PageRank/PageRank.h, PageRank.cu:296-313
 *   mg_stats.edges_traversed   TEPS count                Algorithms/BFS/bfs.cu:483-493
 *   mg_stats.device_ms         "Kernel time took"        enactor_vertex_centric.cuh:2874-2883
 *   mg_options fields          Config parameters         Algorithms/BFS/register.h:27-40,
 *                                                        Algorithms/PageRank/register.h:27-41
 *   mg_csr_from_edges    CsrGraph::FromCoo               b40c/graph/csr_graph.cuh:74-164
 *   mg_mtx_*             ReadMarketStream                b40c/graph/builder/market.cuh:42-166
 *   mg_pick_sources      random source rule              Algorithms/BFS/bfs.cu:340-361
 */
#ifndef MAPGRAPH_C_H_
#define MAPGRAPH_C_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
#define MG_OK                 0
#define MG_ERR_INVALID_ARG    1   /* bad pointer, size, vertex id or option            */
#define MG_ERR_OOM            2   /* device or host allocation failed                   */
#define MG_ERR_DEVICE         3   /* a HIP runtime call or kernel failed                */
#define MG_ERR_OVERFLOW       4   /* a frontier queue overflowed and could not regrow   */
#define MG_ERR_IO             5   /* file could not be opened / read                    */
#define MG_ERR_FORMAT         6   /* malformed Matrix Market input                      */
#define MG_ERR_UNSUPPORTED    7   /* the graph lacks what the call needs (e.g. weights) */
#define MG_ERR_NO_DEVICE      8   /* no HIP device is visible                           */
#define MG_ERR_TIMEOUT        9   /* mg_server_wait: the query did not finish in time   */

/* Copies the last error message of the calling thread (NUL-terminated, truncated to
 * len) and returns its full length. */
size_t mg_last_error(char *buf, size_t len);

/* Library version as "major.minor.patch". */
const char *mg_version(void);

/* ---------------------------------------------------------------- device */
/* Selects the HIP device used by graphs created afterwards on this thread
 * (reference: cudaInit(device), Algorithms/BFS/bfs.cu:145-202). */
int mg_device_init(int device);
/* Number of visible HIP devices (0 when none; never an error). */
int mg_device_count(void);
/* Blocks until all work the library queued on the current device has finished. */
int mg_device_synchronize(void);

/* ---------------------------------------------------------------- graph store */
/* flags for mg_graph_create */
#define MG_GRAPH_WEIGHTED   0x1u  /* edge_values is given (int32 per CSR edge)          */
#define MG_GRAPH_BUILD_CSC  0x2u  /* build the in-edge (CSC) view on the device         */
#define MG_GRAPH_SYMMETRIC  0x4u  /* caller asserts the CSR is symmetric: in-edges of v
                                     equal out-edges of v, with w(u,v) == w(v,u) when
                                     weighted (lets pull steps read the CSR) */
#define MG_GRAPH_RELABEL    0x8u  /* also keep a copy renamed in out-degree order (hubs
                                     first, rows sorted); mg_bfs, mg_sssp and mg_pagerank
                                     run on it (mg_options.layout) and map the source in
                                     and the result back: BFS / SSSP outputs are unchanged,
                                     PageRank sums its in-edges in the copy's order (same
                                     tolerance).  Packs the values the hubs' neighbours
                                     gather into few cache lines; costs a second copy of
                                     the topology. */

typedef struct mg_graph mg_graph;

/* Copies a host CSR graph into device memory (HBM).  row_offsets has n+1 entries,
 * col_idx (and edge_values when MG_GRAPH_WEIGHTED) m entries.  Offsets must be
 * non-decreasing, start at 0, end at m, and every column index must lie in [0, n).
 * The caller's buffers are not retained. */
int mg_graph_create(const int32_t *row_offsets, const int32_t *col_idx,
                    const int32_t *edge_values, int32_t n, int64_t m,
                    uint32_t flags, mg_graph **out);
int mg_graph_destroy(mg_graph *g);
/* The disjoint union of ngraphs resident graphs (of one device) as a new graph, built
 * on the device: part i's vertices v become vertex_offsets[i] + v (vertex_offsets, if
 * not NULL, gets ngraphs + 1 entries, the last one the union's n).  The union is
 * weighted (symmetric) when every part is; it has no CSC view.  The parts stay
 * independent graphs.  No counterpart in the reference (one graph per CsrProblem). */
int mg_graph_union(int32_t ngraphs, mg_graph *const *parts, mg_graph **out, int32_t *vertex_offsets);
/* n, m and the flags a graph was created with. */
int mg_graph_info(const mg_graph *g, int32_t *n, int64_t *m, uint32_t *flags);
/* Device bytes held by the graph (topology + pooled per-search state). */
int mg_graph_device_bytes(const mg_graph *g, int64_t *bytes);
/* Copies the device CSC view back (for tests of the device CSC build).  col_offsets
 * has n+1 entries, row_idx and csr_edge_id m entries; csr_edge_id[k] is the CSR edge
 * id of CSC slot k (reference d_edgeCSC_indices, util.cuh:85-126). */
int mg_graph_get_csc(const mg_graph *g, int32_t *col_offsets, int32_t *row_idx,
                     int32_t *csr_edge_id);

/* ---------------------------------------------------------------- options/stats */
#define MG_ENGINE_AUTO   0   /* specialised kernels where the program allows it         */
#define MG_ENGINE_GAS    1   /* the generic GAS enactor driving the Program functors    */

#define MG_PR_GAS        0   /* reference PageRank program: frontier-activated Jacobi,
                                a vertex stays active while its rank moves by >= 0.01   */
#define MG_PR_ALL        1   /* all-vertex Jacobi for a fixed number of iterations,
                                stopping early once the L1 change falls below tol       */

typedef struct mg_options {
    int32_t iter_num;       /* max iterations; reference "iter_num" (default INT32_MAX)    */
    int32_t threshold;      /* reference "threshold": frontier size above which the
                               two-phase / pull path is taken (default 10000)             */
    int32_t directed;       /* reference "directed": 1 = out-edges only, 0 = out- and
                               in-edges (needs MG_GRAPH_BUILD_CSC or MG_GRAPH_SYMMETRIC)   */
    int32_t engine;         /* MG_ENGINE_AUTO or MG_ENGINE_GAS                             */
    int32_t verbose;        /* reference "verbose": per-iteration frontier trace to stderr */
    int32_t out_on_device;  /* 1: output pointers are device pointers                      */
    int32_t pr_mode;        /* MG_PR_GAS or MG_PR_ALL                                      */
    float   pr_teleport;    /* teleport term t (reference 0.15f); the damping is computed
                               as (1.0f - t) in float, exactly as PageRank.h:269 writes it */
    float   pr_tol;         /* MG_PR_ALL: stop when sum |r_new - r| < pr_tol * n           */
    float   pr_active_eps;  /* MG_PR_GAS activation threshold (reference 0.01f)           */
    double  max_queue_sizing; /* reference "max_queue_sizing" (queue = sizing * n)         */
    int32_t alpha;          /* BFS top-down -> bottom-up switch (frontier edges > unexplored/alpha) */
    int32_t beta;           /* BFS bottom-up -> top-down switch (frontier < n/beta)        */
    int32_t batch_alpha;    /* mg_bfs_batch push -> pull switch (frontier edges > m/batch_alpha); 0 = 4 */
    int32_t batch_beta;     /* mg_bfs_batch pull -> push switch (frontier edges < m/batch_beta); 0 = 16  */
    int32_t sssp_delta;     /* SSSP bucket width: relax only distances below the current
                               bucket bound, raised by sssp_delta when no such distance is
                               left.  mg_sssp: > 0 buckets every level (whole-GPU levels
                               push and carry waiting vertices along); 0 = auto buckets the
                               single-workgroup levels at 4 x the largest weight and runs the
                               whole-GPU levels unbucketed (dense levels pull); -1 =
                               unbucketed.  Batched searches: the width of every level, 0 =
                               about 4 x the mean weight, -1 = unbucketed                  */
    int32_t ppr_iters;      /* personalized PageRank iterations (0 = 20)                   */
    int32_t sssp_bucket_edges; /* mg_sssp: once the pending vertices of a whole-GPU level
                               have fewer out-edges than this, the bucket bound is lifted
                               and the rest runs as Bellman-Ford; 0 = auto (m / 8), -1 =
                               never lifted                                                */
    int32_t level_grid;     /* mg_bfs / mg_sssp: workgroups of the whole-GPU level kernels
                               (1..65536; 0 = auto from n and m)                           */
    int32_t layout;         /* graphs created with MG_GRAPH_RELABEL: 0 = mg_bfs / mg_sssp /
                               mg_pagerank run on the degree-ordered copy, 1 = on the
                               caller's numbering                                          */
    int32_t reserved[1];
} mg_options;

/* Fills *o with the reference defaults. */
void mg_options_default(mg_options *o);

typedef struct mg_stats {
    int64_t edges_traversed;  /* reference TEPS count: sum of out-degrees of the reached
                                 vertices (Algorithms/BFS/bfs.cu:483-493)                  */
    int64_t vertices_reached;
    int64_t edges_examined;   /* edges the kernels actually read (diagnostic)              */
    int32_t iterations;       /* reference "Total iteration"                               */
    int32_t overflow_regrows; /* queue regrow-and-replay events                            */
    double  device_ms;        /* hipEvent time around the search on the graph's stream     */
    double  host_ms;          /* wall time of the call                                     */
    int32_t bottom_up_levels; /* BFS levels run bottom-up                                  */
    int32_t reserved[7];
} mg_stats;

/* ---------------------------------------------------------------- searches */
/* BFS depths from src: depth[src] = 0, unreached = -1 (bfs.h:71-78, 241-253). */
int mg_bfs(mg_graph *g, int32_t src, int32_t *depth_out,
           const mg_options *opt, mg_stats *stats);

/* BFS from src followed by predecessor extraction
 * (Algorithms/BFS/predextract.cuh:246-375, called at bfs.cu:459-468):
 *   pred[v] = max { u in in(v) : depth[v] - depth[u] == 1 }, or -1 if none,
 * for every vertex (the source's value is not a parent, as in the reference).
 * Needs MG_GRAPH_BUILD_CSC or MG_GRAPH_SYMMETRIC.  depth_out may be NULL. */
int mg_bfs_pred(mg_graph *g, int32_t src, int32_t *depth_out, int32_t *pred_out,
                const mg_options *opt, mg_stats *stats);

/* SSSP distances from src over int32 edge weights (needs MG_GRAPH_WEIGHTED):
 * dist[src] = 0, unreached = INT32_MAX (sssp.h:42-43). */
int mg_sssp(mg_graph *g, int32_t src, int32_t *dist_out,
            const mg_options *opt, mg_stats *stats);

/* Connected components: label[v] = smallest vertex id of v's (weakly) connected
 * component (cc.h:118-120, 199-213). */
int mg_cc(mg_graph *g, int32_t *label_out, const mg_options *opt, mg_stats *stats);

/* PageRank in the reference scale: r0 = t, r = t + (1.0f - t) * sum_in r[u]/outdeg[u]
 * with t = pr_teleport (PageRank.h:108-116, 233-283); dangling mass is dropped.
 * MG_PR_GAS runs the reference program (frontier-activated, pr_active_eps, until no
 * vertex moves or iter_num); MG_PR_ALL updates every vertex for iter_num
 * iterations or until sum |r_new - r| < pr_tol * n.  Needs MG_GRAPH_BUILD_CSC or
 * MG_GRAPH_SYMMETRIC.  Sums run in a fixed order, so results are deterministic. */
int mg_pagerank(mg_graph *g, float *rank_out, const mg_options *opt, mg_stats *stats);

/* ---------------------------------------------------------------- batched searches */
#define MG_BATCH_MAX 256  /* sources per batched call */

/* BFS from k <= MG_BATCH_MAX sources in one device-driven search: depth_out is
 * k x n, row s holding the depths from srcs[s] (the reference runs sources one
 * after another, Algorithms/BFS/bfs.cu:402-417).  Row s equals mg_bfs(g, srcs[s]).
 * mg_stats.edges_traversed and vertices_reached sum the per-source counts.
 * directed=0 needs a graph created with MG_GRAPH_SYMMETRIC.  The dense-level pull
 * direction needs in-edges (symmetric graph or MG_GRAPH_BUILD_CSC); its switch is
 * set by mg_options.batch_alpha / batch_beta. */
int mg_bfs_batch(mg_graph *g, const int32_t *srcs, int32_t k, int32_t *depth_out,
                 const mg_options *opt, mg_stats *stats);

/* SSSP from k <= MG_BATCH_MAX sources in one device-driven search: dist_out is
 * k x n, row s equal to mg_sssp(g, srcs[s]) (unreached = INT32_MAX; the
 * reference runs sources one after another, Algorithms/SSSP/sssp.cu:310-428).
 * Needs MG_GRAPH_WEIGHTED with non-negative weights; directed=0 needs
 * MG_GRAPH_SYMMETRIC.  iter_num bounds the relaxation rounds (rows equal the
 * single-source result when the search converges within it).  mg_stats
 * edges_traversed / vertices_reached sum the per-source counts. */
int mg_sssp_batch(mg_graph *g, const int32_t *srcs, int32_t k, int32_t *dist_out,
                  const mg_options *opt, mg_stats *stats);

/* Batched searches with source sets: npairs (row, source) pairs, row_of[i] in [0, k),
 * every row with at least one source.  Row r of the k x n result holds the depths
 * (distances) from the set of its sources, as a search seeded with several sources
 * (the reference's num_srcs / srcs, Algorithms/BFS/bfs.h:62-126, which runs them as
 * one frontier).  On a disjoint union of graphs (mg_graph_union) a row may carry one
 * source in each part: the parts' searches never meet, so each part's segment of the
 * row equals the single-source search on that part -- several graphs of uneven size
 * advance in one batched search.  Same options and requirements as mg_bfs_batch /
 * mg_sssp_batch; mg_stats sums the rows' counts. */
int mg_bfs_batch_sets(mg_graph *g, const int32_t *row_of, const int32_t *srcs, int32_t npairs, int32_t k,
                      int32_t *depth_out, const mg_options *opt, mg_stats *stats);
int mg_sssp_batch_sets(mg_graph *g, const int32_t *row_of, const int32_t *srcs, int32_t npairs, int32_t k,
                       int32_t *dist_out, const mg_options *opt, mg_stats *stats);

#define MG_PPR_MAX 64  /* sources per batched personalized PageRank call */

/* Personalized PageRank from k <= MG_PPR_MAX sources, FP32 (new here; the
 * reference has only global PageRank, Algorithms/PageRank/PageRank.h:233-283,
 * whose conventions it keeps): x_0 = e_s and
 *   x_{i+1}[u] = (1 - t) * sum_{v -> u} x_i[v] / outdeg(v)  (+ t at u = s),
 * t = opt->pr_teleport (default 0.15), opt->ppr_iters iterations (default 20),
 * in-edges of a symmetric graph or the CSC.  rank_out is k x n.  Each row's
 * floating-point operations are fixed by the graph, so a row is bitwise the same
 * whatever other sources share the call. */
int mg_ppr_batch(mg_graph *g, const int32_t *srcs, int32_t k, float *rank_out,
                 const mg_options *opt, mg_stats *stats);

/* ---------------------------------------------------------------- autotuning */
/* Picks, for one graph and one search kind, the mg_options fields that decide how the
 * search maps onto the GPU (no counterpart in the reference, which fixes its kernel
 * policies per SM version at compile time, GASengine/enactor_vertex_centric.cuh:
 * 2910-2966, and takes one `threshold`, Algorithms/BFS/register.h:39):
 *   MG_TUNE_BFS         alpha, beta, level_grid            (mg_bfs)
 *   MG_TUNE_SSSP        sssp_delta, sssp_bucket_edges, level_grid   (mg_sssp)
 *   MG_TUNE_BFS_BATCH   batch_alpha, batch_beta             (mg_bfs_batch)
 *   MG_TUNE_SSSP_BATCH  sssp_delta                          (mg_sssp_batch)
 * Coordinate descent from *start (NULL: the defaults) over candidate values, two
 * passes: each setting is timed on the sample sources (device time, median of reps
 * runs after a warm-up; single-source kinds: up to 64 sources, summed; batched kinds:
 * one call with all of them, up to MG_BATCH_MAX) and kept when at least 3% faster.
 * Every candidate's results must equal the starting setting's bit for bit, else
 * MG_ERR_DEVICE.  *best receives the chosen options (the other fields as *start). */
#define MG_TUNE_BFS        0
#define MG_TUNE_SSSP       1
#define MG_TUNE_BFS_BATCH  2
#define MG_TUNE_SSSP_BATCH 3

typedef struct mg_tune_report {
    double  start_ms;         /* device ms of the starting setting (as timed by the tuner) */
    double  best_ms;          /* device ms of the chosen setting                           */
    int32_t candidates;       /* settings timed                                            */
    int32_t reserved[7];
} mg_tune_report;

int mg_autotune(mg_graph *g, int32_t what, const int32_t *srcs, int32_t nsrc, int32_t reps,
                const mg_options *start, mg_options *best, mg_tune_report *report);

/* ---------------------------------------------------------------- serving */
/* A request loop over resident graphs (new here: the reference runs one search at a
 * time, e.g. Algorithms/BFS/bfs.cu:402-417).  Queries are served with continuous
 * batching: each (graph, kind) engine keeps one batched search running and seeds
 * newly admitted queries into free slots between chunks of levels, so queries that
 * arrive while others are on the GPU join them; engines of different graphs run on
 * their own streams and overlap.  Engine state is allocated within hbm_budget; a
 * query waits while no slot or memory is free.  A query's result equals the
 * single-query result: BFS depths (mg_bfs), SSSP distances (mg_sssp), personalized
 * PageRank ranks (mg_ppr_batch, bitwise).  Graphs must outlive the server. */
typedef struct mg_server mg_server;

#define MG_Q_BFS  0
#define MG_Q_SSSP 1
#define MG_Q_PPR  2

typedef struct mg_server_config {
    int64_t hbm_budget;       /* device bytes the engines may hold (default 64 GiB)      */
    int32_t max_slots;        /* BFS / SSSP slots per engine, <= MG_BATCH_MAX (256)      */
    int32_t ppr_iters;        /* personalized PageRank iterations (default 20)           */
    float   ppr_teleport;     /* personalized PageRank teleport t (default 0.15)         */
    int32_t steps_per_round;  /* levels per engine round between admissions (default 8)  */
    int32_t sssp_delta;       /* SSSP bucket width, as mg_options.sssp_delta (0 = auto)  */
    int32_t thin_levels;      /* BFS / SSSP on a graph whose probe BFS (from a sampled
                                 vertex, at mg_server_add_graph) takes at least this many
                                 levels run as thin searches: one workgroup runs a whole
                                 query inside one launch (0 = 256; -1 = never)          */
    int32_t thin_lanes;       /* concurrent thin queries per (graph, kind) (0 = 64)      */
    int32_t thin_streams;     /* thin launches in flight per (graph, kind), 1..8 (0 = 1) */
    int32_t thin_slice_us;    /* a thin search pauses at a level boundary after this many
                                 microseconds in a launch and continues in a later one
                                 (0 = 5000; -1 = runs to the end in one launch)          */
    int32_t ppr_gate_us;      /* personalized PageRank cohorts: an idle engine starts when
                                 ppr_batch_min queries wait or the oldest has waited this
                                 long (0 = 20000; -1 = every query starts on arrival)   */
    int32_t ppr_batch_min;    /* (0 = 8, at most 64)                                     */
    int32_t plain_kinds;      /* bit k set (1 << MG_Q_BFS, 1 << MG_Q_SSSP): those queries on
                                 graphs that are not thin run as whole-GPU searches
                                 (mg_bfs / mg_sssp, the graph's twin included), two at a
                                 time per (graph, kind), each worker on its own view of the
                                 graph and dedicated stream, instead of in batches
                                 (default 1 << MG_Q_SSSP; 0 = all batched)               */
} mg_server_config;

typedef struct mg_query_info {
    int32_t status;           /* MG_OK or the error that failed the query                */
    int32_t kind, graph, src;
    int64_t edges_traversed;  /* TEPS count (PageRank: m x iterations)                   */
    int64_t vertices_reached;
    double  t_submit_ms, t_admit_ms, t_done_ms;  /* server clock (mg_server_now_ms)     */
    int32_t reserved[4];
} mg_query_info;

typedef struct mg_server_stats {
    int64_t submitted, completed, rounds, edges_traversed, device_bytes;
    double  latency_ms_sum;
    int32_t thin_engines;     /* (graph, kind) engines serving thin searches              */
    int32_t pad;
    double  dispatch_busy_ms; /* dispatcher time spent advancing engines (host side)     */
    int64_t dispatch_passes;  /* dispatcher passes over the engines                      */
    int32_t reserved[2];
} mg_server_stats;

void mg_server_config_default(mg_server_config *c);
int mg_server_create(const mg_server_config *cfg, mg_server **out);
/* Registers a graph; queries name it by *graph_id. */
int mg_server_add_graph(mg_server *s, mg_graph *g, int32_t *graph_id);
/* Queues a query.  out receives n int32 (BFS depths / SSSP distances) or n float
 * (PageRank) when the query finishes; host memory from mg_host_alloc keeps the copy
 * asynchronous (other host memory works but blocks the dispatcher while it copies);
 * out_on_device = 1 for a device pointer; NULL discards the result. */
int mg_server_submit(mg_server *s, int32_t graph_id, int32_t kind, int32_t src, void *out,
                     int32_t out_on_device, int64_t *query_id);
/* Waits for a query (timeout_ms < 0: forever), fills *info and forgets the query. */
int mg_server_wait(mg_server *s, int64_t query_id, int32_t timeout_ms, mg_query_info *info);
int mg_server_now_ms(mg_server *s, double *ms);
int mg_server_get_stats(mg_server *s, mg_server_stats *st);
/* Stops the dispatcher after the round in flight and frees all engine state;
 * unfinished queries are dropped. */
int mg_server_destroy(mg_server *s);
/* Pinned host memory for results. */
void *mg_host_alloc(size_t bytes);
void mg_host_free(void *p);

/* ---------------------------------------------------------------- host graph tools */
/* These run on the host only and need no device. */

/* R-MAT edge count ef * 2^scale. */
int64_t mg_rmat_num_edges(int32_t scale, int32_t edge_factor);
/* Seeded R-MAT edges: edge k picks, for each of scale levels, one quadrant with
 * probabilities (a, b, c, 1-a-b-c) from a counter-based hash of (seed, k, level);
 * no noise, no vertex permutation.  Writes mg_rmat_num_edges() pairs. */
int mg_gen_rmat_edges(int32_t scale, int32_t edge_factor, double a, double b, double c,
                      uint64_t seed, int32_t *src, int32_t *dst);
/* Undirected width x height 4-neighbour grid (vertex id = row * width + col) with
 * exactly floor(delete_frac * E) edges removed by a seeded hash order.  Writes the
 * kept undirected edges (u < v) and their count; the buffers need room for
 * mg_grid_num_edges(). */
int64_t mg_grid_num_edges(int32_t width, int32_t height);
int mg_gen_grid_edges(int32_t width, int32_t height, double delete_frac, uint64_t seed,
                      int32_t *src, int32_t *dst, int64_t *m_out);

/* flags for mg_csr_from_edges */
#define MG_CSR_SYMMETRIZE   0x1u  /* add (v,u) for every (u,v)                      */
#define MG_CSR_DROP_LOOPS   0x2u  /* drop (u,u)                                     */
#define MG_CSR_DEDUPE       0x4u  /* keep one copy of repeated (u,v)                */
#define MG_CSR_SORT_COLS    0x8u  /* sort each row's columns ascending; without it
                                     rows keep input order (stable sort by row, as
                                     CsrGraph::FromCoo, csr_graph.cuh:74-122)      */
/* Builds a CSR from an edge list.  row_offsets needs n+1 entries; col_idx needs
 * capacity entries (2*m_in suffices with MG_CSR_SYMMETRIZE, m_in otherwise).
 * When vals/vals_out are non-NULL, each edge's value travels with it (the
 * reverse edge of a symmetrized pair gets the same value). */
int mg_csr_from_edges(int32_t n, int64_t m_in, const int32_t *src, const int32_t *dst,
                      const int32_t *vals, uint32_t flags, int32_t *row_offsets,
                      int32_t *col_idx, int32_t *vals_out, int64_t capacity,
                      int64_t *m_out);
/* Symmetric hashed edge weights: w = 1 + hash(seed, min(u,v), max(u,v)) mod wmax. */
int mg_edge_weights_hash(int32_t n, const int32_t *row_offsets, const int32_t *col_idx,
                         int64_t m, uint64_t seed, int32_t wmax, int32_t *w_out);
/* Host CSC (in-edge) view by a stable counting sort on the column. */
int mg_csc_from_csr(int32_t n, int64_t m, const int32_t *row_offsets,
                    const int32_t *col_idx, int32_t *col_offsets, int32_t *row_idx,
                    int32_t *csr_edge_id);
/* Degree-order relabelling: new_of_old[v] = rank of v by out-degree (highest first,
 * ties by id); the renamed CSR lists every row's renamed targets ascending, values
 * travelling with their edges (vals / vals_out may be NULL).  No reference
 * counterpart: a layout transform for locality (the hubs' entries pack together). */
int mg_csr_relabel_degree(int32_t n, int64_t m, const int32_t *row_offsets, const int32_t *col_idx,
                          const int32_t *vals, int32_t *new_of_old, int32_t *row_offsets_out,
                          int32_t *col_idx_out, int32_t *vals_out);
/* k seeded uniform draws (with replacement) among vertices of out-degree > 0. */
int mg_pick_sources(int32_t n, const int32_t *row_offsets, int32_t k, uint64_t seed,
                    int32_t *out);
/* Matrix Market: header pass, then the (0-based) entries.  Values default to 1
 * when the file has no value column; values are truncated to int32 as the
 * reference's (Value) cast does (market.cuh:116-121). */
int mg_mtx_read_header(const char *path, int32_t *n, int64_t *nnz, int32_t *symmetric);
int mg_mtx_read(const char *path, int32_t *src, int32_t *dst, int32_t *val,
                int64_t capacity, int64_t *nnz_out);

#ifdef __cplusplus
}
#endif

#endif /* MAPGRAPH_C_H_ */
