// pagerank -- PageRank as a GAS program (port of Algorithms/PageRank/PageRank.h:33-409):
// gather rank[u] / outdeg[u] over in-edges, apply 0.15 + 0.85 * sum, a vertex stays
// active while its rank moves by >= 0.01; duplicates are not allowed.
#pragma once

#include <cmath>

#include "../gas.h"

struct pagerank {
    typedef float DataType;
    typedef DataType MiscType;
    typedef float GatherType;
    typedef int VertexId;
    typedef int SizeT;

    static constexpr DataType INIT_VALUE = 0.0f;
    static constexpr bool allow_duplicates = false;

    struct VertexType {
        int nodes = 0;
        int edges = 0;
        DataType *d_dists = nullptr;     // the rank
        int *d_num_out_edge = nullptr;   // out-degree
    };

    struct EdgeType {
        int nodes = 0;
        int edges = 0;
    };

    // PageRank.h:73-190: rank = 0.15, frontier = all vertices, out-degree from the
    // row offsets (plus in-degree when the graph is treated as undirected)
    static void Initialize(const int directed, const int nodes, const int edges, int num_srcs, int *srcs,
                           int *d_row_offsets, int *d_column_indices, int *d_column_offsets, int *d_row_indices,
                           DataType *d_edge_values, VertexType &vertex_list, EdgeType &edge_list,
                           int *d_frontier_keys[3], MiscType *d_frontier_values[3]) {
        (void)num_srcs, (void)srcs, (void)d_column_indices, (void)d_row_indices, (void)d_edge_values,
            (void)d_frontier_values;
        vertex_list.nodes = edge_list.nodes = nodes;
        vertex_list.edges = edge_list.edges = edges;
        if (!vertex_list.d_dists) (void)mapgraph::device_alloc(&vertex_list.d_dists, nodes);
        if (!vertex_list.d_num_out_edge) (void)mapgraph::device_alloc(&vertex_list.d_num_out_edge, nodes);
        (void)mapgraph::fill(vertex_list.d_dists, 0.15f, nodes);
        (void)mapgraph::iota(d_frontier_keys[0], nodes);
        (void)mapgraph::iota(d_frontier_keys[1], nodes);
        (void)mapgraph::degrees(d_row_offsets, directed == 0 ? d_column_offsets : nullptr, nodes,
                                vertex_list.d_num_out_edge);
    }

    static void Release(VertexType &vertex_list, EdgeType &) {
        if (vertex_list.d_dists) (void)hipFree(vertex_list.d_dists);
        if (vertex_list.d_num_out_edge) (void)hipFree(vertex_list.d_num_out_edge);
        vertex_list.d_dists = nullptr;
        vertex_list.d_num_out_edge = nullptr;
    }

    static SrcVertex srcVertex() { return ALL; }
    static GatherEdges gatherOverEdges() { return GATHER_IN_EDGES; }
    static ApplyVertices applyOverEdges() { return APPLY_FRONTIER; }
    static PostApplyVertices postApplyOverEdges() { return POST_APPLY_FRONTIER; }
    static ExpandEdges expandOverEdges() { return EXPAND_OUT_EDGES; }

    // PageRank.h:233-243
    struct gather_edge {
        __device__ void operator()(const int, const int, const int neighbor_id_in, VertexType &vertex_list,
                                   EdgeType &, GatherType &new_value) {
            const DataType nb_dist = vertex_list.d_dists[neighbor_id_in];
            new_value = nb_dist / (DataType)vertex_list.d_num_out_edge[neighbor_id_in];
        }
    };

    struct gather_sum {
        __device__ GatherType operator()(const GatherType &left, const GatherType &right) { return left + right; }
    };

    // PageRank.h:257-283
    struct apply {
        __device__ void operator()(const int vertex_id, const int, GatherType gathervalue, VertexType &vertex_list,
                                   EdgeType &, char &changed) {
            const DataType oldvalue = vertex_list.d_dists[vertex_id];
            const DataType newvalue = 0.15f + (1.0f - 0.15f) * gathervalue;
            changed = fabsf(oldvalue - newvalue) < 0.01f ? 0 : 1;
            vertex_list.d_dists[vertex_id] = newvalue;
        }
    };

    struct post_apply {
        __device__ void operator()(const int, VertexType &, EdgeType &, GatherType *) {}
    };

    struct expand_vertex {
        __device__ bool operator()(const int, const char changed, VertexType &, EdgeType &) { return changed; }
    };

    // PageRank.h:355-362 (the engine sets frontier = -1 before the call)
    struct expand_edge {
        __device__ void operator()(const bool changed, const int, const int, const int neighbor_id_in, const int,
                                   VertexType &, EdgeType &, int &frontier, MiscType &) {
            if (changed) frontier = neighbor_id_in;
        }
    };

    struct contract {
        __device__ void operator()(const int, int &, VertexType &, EdgeType &, GatherType *, MiscType &) {}
    };

    static void extractResult(VertexType &vertex_list, DataType *h_output) {
        (void)hipMemcpy(h_output, vertex_list.d_dists, sizeof(DataType) * vertex_list.nodes, hipMemcpyDeviceToHost);
    }
};
