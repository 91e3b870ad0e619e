// csr_problem.h -- device graph store for a GAS program (the role of
// GASengine/csr_problem.cuh:112-831).
//
// FromHostProblem copies a host CSR (+ its CSC when given) into HBM and derives the
// CSR edge id of every CSC slot by a stable counting sort on the column (what
// util.cuh:85-126 computes with thrust::stable_sort_by_key).  FromDeviceGraph wraps
// arrays that already live in HBM without copying (the C-ABI's mg_graph uses it).
// The store keeps one GraphSlice (single GPU, as the reference does in practice:
// ExtractResults throws for more than one slice, csr_problem.cuh:387-391).
// Frontier storage is O(n) (three key/value queues of n entries for Program::
// Initialize plus the engine's bitmaps), not the reference's O(max_queue_sizing * m).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <vector>

#include "config.h"
#include "gas.h"

namespace GASengine {

namespace detail {
// Program::Release(VertexType&, EdgeType&) when the program defines it.
template <class P, class = void>
struct has_release : std::false_type {};
template <class P>
struct has_release<P, std::void_t<decltype(P::Release(std::declval<typename P::VertexType &>(),
                                                      std::declval<typename P::EdgeType &>()))>>
    : std::true_type {};
}  // namespace detail

template <typename _Program, typename _VertexId, typename _SizeT, typename _EValue, bool MARK_PREDECESSORS,
          bool WITH_VALUE>
struct CsrProblem {
    typedef _Program Program;
    typedef _VertexId VertexId;
    typedef _SizeT SizeT;
    typedef _EValue EValue;
    typedef typename Program::DataType DataType;
    typedef typename Program::MiscType MiscType;
    typedef typename Program::GatherType GatherType;

    struct GraphSlice {
        int gpu = 0;
        hipStream_t stream = nullptr;
        SizeT nodes = 0, edges = 0;
        VertexId *d_column_indices = nullptr;  // CSR
        SizeT *d_row_offsets = nullptr;
        EValue *d_edge_values = nullptr;
        VertexId *d_row_indices = nullptr;     // CSC (nullptr when absent)
        SizeT *d_column_offsets = nullptr;
        SizeT *d_edgeCSC_indices = nullptr;    // CSR edge id of each CSC slot
        bool symmetric = false;                // CSR == CSC (in-edges read from the CSR)
        typename Program::VertexType vertex_list;
        typename Program::EdgeType edge_list;
        struct {
            VertexId *d_keys[3] = {nullptr, nullptr, nullptr};
            MiscType *d_values[3] = {nullptr, nullptr, nullptr};
        } frontier_queues;
        char *d_changed = nullptr;
        bool owns_topology = false;

        ~GraphSlice() {
            if (detail::has_release<Program>::value) release_program(vertex_list, edge_list);
            for (int i = 0; i < 3; ++i) {
                if (frontier_queues.d_keys[i]) (void)hipFree(frontier_queues.d_keys[i]);
                if (frontier_queues.d_values[i]) (void)hipFree(frontier_queues.d_values[i]);
            }
            if (d_changed) (void)hipFree(d_changed);
            if (owns_topology) {
                if (d_column_indices) (void)hipFree(d_column_indices);
                if (d_row_offsets) (void)hipFree(d_row_offsets);
                if (d_edge_values) (void)hipFree(d_edge_values);
                if (d_row_indices) (void)hipFree(d_row_indices);
                if (d_column_offsets) (void)hipFree(d_column_offsets);
                if (d_edgeCSC_indices) (void)hipFree(d_edgeCSC_indices);
            }
        }

      private:
        template <class P = Program>
        static void release_program(typename P::VertexType &v, typename P::EdgeType &e) {
            if constexpr (detail::has_release<P>::value) P::Release(v, e);
        }
    };

    mapgraph::Config cfg;
    SizeT nodes = 0, edges = 0;
    int num_gpus = 1;
    std::vector<GraphSlice *> graph_slices;

    explicit CsrProblem(const mapgraph::Config &c = mapgraph::Config()) : cfg(c) {}
    CsrProblem(const CsrProblem &) = delete;
    CsrProblem &operator=(const CsrProblem &) = delete;
    ~CsrProblem() {
        for (auto *s : graph_slices) delete s;
    }

    // csr_problem.cuh:401-621.  h_row_indices / h_column_offsets (the CSC) may be
    // nullptr, in which case the CSC is derived from the CSR.  Returns hipSuccess or
    // the first HIP error (hipErrorInvalidValue for more than one GPU).
    hipError_t FromHostProblem(bool stream_from_host, SizeT nodes_, SizeT edges_, VertexId *h_column_indices,
                               SizeT *h_row_offsets, EValue *h_edge_values, VertexId *h_row_indices,
                               SizeT *h_column_offsets, int num_gpus_, int directed) {
        (void)stream_from_host, (void)directed;
        if (num_gpus_ != 1) return hipErrorInvalidValue;
        clear();
        nodes = nodes_;
        edges = edges_;
        num_gpus = 1;
        GraphSlice *s = new GraphSlice();
        graph_slices.push_back(s);
        s->gpu = cfg.getParameter<int>("device");
        hipError_t e;
        if ((e = hipSetDevice(s->gpu)) != hipSuccess) return e;
        s->nodes = nodes;
        s->edges = edges;
        s->owns_topology = true;
        // CSC by a stable counting sort on the column: slot k holds CSR edge id
        std::vector<SizeT> co(nodes + 1, 0), eid(edges);
        std::vector<VertexId> ri(edges);
        for (SizeT k = 0; k < edges; ++k) co[h_column_indices[k] + 1]++;
        for (SizeT v = 0; v < nodes; ++v) co[v + 1] += co[v];
        {
            std::vector<SizeT> fill(co.begin(), co.end() - 1);
            for (SizeT u = 0; u < nodes; ++u)
                for (SizeT k = h_row_offsets[u]; k < h_row_offsets[u + 1]; ++k) {
                    const SizeT slot = fill[h_column_indices[k]]++;
                    ri[slot] = u;
                    eid[slot] = k;
                }
        }
        if (h_column_offsets && h_row_indices) {  // a caller-supplied CSC must agree
            for (SizeT v = 0; v <= nodes; ++v)
                if (h_column_offsets[v] != co[v]) return hipErrorInvalidValue;
        }
        if ((e = upload(&s->d_row_offsets, h_row_offsets, nodes + 1)) != hipSuccess ||
            (e = upload(&s->d_column_indices, h_column_indices, edges)) != hipSuccess ||
            (e = upload(&s->d_column_offsets, co.data(), nodes + 1)) != hipSuccess ||
            (e = upload(&s->d_row_indices, ri.data(), edges)) != hipSuccess ||
            (e = upload(&s->d_edgeCSC_indices, eid.data(), edges)) != hipSuccess)
            return e;
        if (WITH_VALUE && h_edge_values) {
            if ((e = upload(&s->d_edge_values, h_edge_values, edges)) != hipSuccess) return e;
        }
        return allocate_frontier(s);
    }

    // Wraps device arrays owned by the caller (no copy).
    hipError_t FromDeviceGraph(int device, SizeT nodes_, SizeT edges_, VertexId *d_col, SizeT *d_row_off,
                               EValue *d_vals, SizeT *d_col_off, VertexId *d_row_idx, SizeT *d_csc_eid,
                               bool symmetric, hipStream_t stream) {
        clear();
        nodes = nodes_;
        edges = edges_;
        GraphSlice *s = new GraphSlice();
        graph_slices.push_back(s);
        s->gpu = device;
        s->stream = stream;
        s->nodes = nodes;
        s->edges = edges;
        s->d_column_indices = d_col;
        s->d_row_offsets = d_row_off;
        s->d_edge_values = d_vals;
        s->d_column_offsets = d_col_off;
        s->d_row_indices = d_row_idx;
        s->d_edgeCSC_indices = d_csc_eid;
        s->symmetric = symmetric;
        s->owns_topology = false;
        return allocate_frontier(s);
    }

    // csr_problem.cuh:371-396
    hipError_t ExtractResults(DataType *h_values) {
        if (graph_slices.size() != 1) return hipErrorInvalidValue;
        hipError_t e = hipSetDevice(graph_slices[0]->gpu);
        if (e != hipSuccess) return e;
        Program::extractResult(graph_slices[0]->vertex_list, h_values);
        return hipGetLastError();
    }

  private:
    void clear() {
        for (auto *s : graph_slices) delete s;
        graph_slices.clear();
    }

    template <class T>
    static hipError_t upload(T **d, const T *h, size_t count) {
        hipError_t e = hipMalloc((void **)d, (count ? count : 1) * sizeof(T));
        if (e != hipSuccess || count == 0) return e;
        return hipMemcpy(*d, h, count * sizeof(T), hipMemcpyHostToDevice);
    }

    static hipError_t allocate_frontier(GraphSlice *s) {
        hipError_t e;
        const size_t n = s->nodes > 0 ? (size_t)s->nodes : 1;
        for (int i = 0; i < 3; ++i) {
            if ((e = hipMalloc((void **)&s->frontier_queues.d_keys[i], n * sizeof(VertexId))) != hipSuccess) return e;
            if ((e = hipMalloc((void **)&s->frontier_queues.d_values[i], n * sizeof(MiscType))) != hipSuccess)
                return e;
        }
        return hipMalloc((void **)&s->d_changed, n);
    }
};

}  // namespace GASengine
