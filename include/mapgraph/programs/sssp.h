// sssp -- single-source shortest paths as a GAS program (port of
// Algorithms/SSSP/sssp.h:33-409): push-style frontier Bellman-Ford over int32
// weights, unreached = INT_MAX.  Device functors keep the reference semantics.
#pragma once

#include <vector>

#include "../gas.h"

struct sssp {
    typedef int DataType;
    static constexpr DataType INIT_VALUE = INT_MAX;
    typedef DataType MiscType;
    typedef DataType GatherType;
    typedef int VertexId;
    typedef int SizeT;

    static constexpr bool allow_duplicates = true;

    struct VertexType {
        int nodes = 0;
        int edges = 0;
        DataType *d_dists_out = nullptr;  // new value computed by apply
        DataType *d_dists = nullptr;      // the distance committed by post_apply
    };

    struct EdgeType {
        int nodes = 0;
        int edges = 0;
        DataType *d_weights = nullptr;
    };

    // sssp.h:83-176: dists = INT_MAX, sources = 0, dists_out = dists, frontier =
    // sources.  d_weights points at the graph's edge values (the reference copies
    // them, sssp.h:115-126; they are never written).
    static void Initialize(const int directed, const int nodes, const int edges, int num_srcs, int *srcs,
                           int *d_row_offsets, int *d_column_indices, int *d_column_offsets, int *d_row_indices,
                           DataType *d_edge_values, VertexType &vertex_list, EdgeType &edge_list,
                           int *d_frontier_keys[3], MiscType *d_frontier_values[3]) {
        (void)directed, (void)d_row_offsets, (void)d_column_indices, (void)d_column_offsets, (void)d_row_indices;
        vertex_list.nodes = edge_list.nodes = nodes;
        vertex_list.edges = edge_list.edges = edges;
        if (!vertex_list.d_dists) (void)mapgraph::device_alloc(&vertex_list.d_dists, nodes);
        if (!vertex_list.d_dists_out) (void)mapgraph::device_alloc(&vertex_list.d_dists_out, nodes);
        edge_list.d_weights = d_edge_values;
        (void)mapgraph::fill(vertex_list.d_dists, INIT_VALUE, nodes);
        const int zero = 0;
        for (int i = 0; i < num_srcs; ++i)
            (void)hipMemcpy(vertex_list.d_dists + srcs[i], &zero, sizeof zero, hipMemcpyHostToDevice);
        (void)hipMemcpy(vertex_list.d_dists_out, vertex_list.d_dists, nodes * sizeof(DataType), hipMemcpyDeviceToDevice);
        (void)hipMemcpy(d_frontier_keys[0], srcs, num_srcs * sizeof(int), hipMemcpyHostToDevice);
        (void)hipMemcpy(d_frontier_keys[1], srcs, num_srcs * sizeof(int), hipMemcpyHostToDevice);
        std::vector<MiscType> zeros(num_srcs, 0);
        (void)hipMemcpy(d_frontier_values[0], zeros.data(), num_srcs * sizeof(MiscType), hipMemcpyHostToDevice);
        (void)hipMemcpy(d_frontier_values[1], zeros.data(), num_srcs * sizeof(MiscType), hipMemcpyHostToDevice);
    }

    static void Release(VertexType &vertex_list, EdgeType &edge_list) {
        if (vertex_list.d_dists) (void)hipFree(vertex_list.d_dists);
        if (vertex_list.d_dists_out) (void)hipFree(vertex_list.d_dists_out);
        vertex_list.d_dists = vertex_list.d_dists_out = nullptr;
        edge_list.d_weights = nullptr;
    }

    static SrcVertex srcVertex() { return SINGLE; }
    static GatherEdges gatherOverEdges() { return NO_GATHER_EDGES; }
    static ApplyVertices applyOverEdges() { return APPLY_FRONTIER; }
    static PostApplyVertices postApplyOverEdges() { return POST_APPLY_FRONTIER; }
    static ExpandEdges expandOverEdges() { return EXPAND_OUT_EDGES; }

    struct gather_vertex {
        __device__ void operator()(const int, const GatherType, VertexType &, EdgeType &) {}
    };

    struct gather_edge {
        __device__ void operator()(const int, const int, const int, VertexType &, EdgeType &, GatherType &) {}
    };

    struct gather_sum {
        __device__ GatherType operator()(const GatherType &left, const GatherType &right) { return left + right; }
    };

    // sssp.h:242-269: keep the smaller of the committed distance and the gathered
    // minimum; changed on iteration 0 or when it moved
    struct apply {
        __device__ void operator()(const VertexId vertex_id, const int iteration, GatherType gathervalue,
                                   VertexType &vertex_list, EdgeType &, char &changed) {
            const DataType oldvalue = vertex_list.d_dists[vertex_id];
            const DataType newvalue = min(oldvalue, gathervalue);
            changed = (iteration == 0 || oldvalue != newvalue) ? 1 : 0;
            vertex_list.d_dists_out[vertex_id] = newvalue;
        }
    };

    // sssp.h:272-280: commit and reset the gather slot
    struct post_apply {
        __device__ void operator()(const VertexId vertex_id, VertexType &vertex_list, EdgeType &,
                                   GatherType *gather_tmp) {
            vertex_list.d_dists[vertex_id] = vertex_list.d_dists_out[vertex_id];
            gather_tmp[vertex_id] = INIT_VALUE;
        }
    };

    struct expand_vertex {
        __device__ bool operator()(const VertexId, const char changed, VertexType &, EdgeType &) {
            return changed == 1;
        }
    };

    // sssp.h:342-357: offer src_dist + w when it improves on the neighbour
    struct expand_edge {
        __device__ void operator()(const bool, const int, const VertexId vertex_id, const VertexId neighbor_id_in,
                                   const VertexId edge_id, VertexType &vertex_list, EdgeType &edge_list,
                                   int &frontier, MiscType &misc_value) {
            const int src_dist = vertex_list.d_dists[vertex_id];
            const int dst_dist = vertex_list.d_dists[neighbor_id_in];
            const DataType edge_value = edge_list.d_weights[edge_id];
            frontier = (dst_dist > src_dist + edge_value) ? neighbor_id_in : -1;
            misc_value = src_dist + edge_value;
        }
    };

    // sssp.h:384-401: minimum offer per vertex
    struct contract {
        __device__ void operator()(const int, int &vertex_id, VertexType &, EdgeType &, GatherType *gather_tmp,
                                   MiscType &misc_value) {
            atomicMin(&gather_tmp[vertex_id], misc_value);
        }
    };

    static void extractResult(VertexType &vertex_list, DataType *h_output) {
        (void)hipMemcpy(h_output, vertex_list.d_dists, sizeof(DataType) * vertex_list.nodes, hipMemcpyDeviceToHost);
    }
};
