// cc -- connected components by min-label propagation as a GAS program (port of
// Algorithms/CC/cc.h:32-349).  label[v] starts at v; every vertex starts in the
// frontier; labels shrink to the component minimum.
#pragma once

#include "../gas.h"

struct cc {
    typedef int DataType;
    typedef DataType MiscType;
    typedef DataType GatherType;
    typedef int VertexId;
    typedef int SizeT;

    static constexpr int INIT_VALUE = INT_MAX;
    static constexpr bool allow_duplicates = true;

    struct VertexType {
        int nodes = 0;
        int edges = 0;
        DataType *d_dists_out = nullptr;  // new value computed by apply
        DataType *d_dists = nullptr;      // the label committed by post_apply
    };

    struct EdgeType {
        int nodes = 0;
        int edges = 0;
    };

    static SrcVertex srcVertex() { return ALL; }
    static GatherEdges gatherOverEdges() { return NO_GATHER_EDGES; }
    static ApplyVertices applyOverEdges() { return APPLY_FRONTIER; }
    static PostApplyVertices postApplyOverEdges() { return POST_APPLY_FRONTIER; }
    static ExpandEdges expandOverEdges() { return EXPAND_OUT_EDGES; }

    // cc.h:93-148: label = vertex id, frontier = all vertices
    static void Initialize(const int directed, const int nodes, const int edges, int num_srcs, int *srcs,
                           int *d_row_offsets, int *d_column_indices, int *d_column_offsets, int *d_row_indices,
                           int *d_edge_values, VertexType &vertex_list, EdgeType &edge_list,
                           int *d_frontier_keys[3], MiscType *d_frontier_values[3]) {
        (void)directed, (void)num_srcs, (void)srcs, (void)d_row_offsets, (void)d_column_indices,
            (void)d_column_offsets, (void)d_row_indices, (void)d_edge_values;
        vertex_list.nodes = edge_list.nodes = nodes;
        vertex_list.edges = edge_list.edges = edges;
        if (!vertex_list.d_dists) (void)mapgraph::device_alloc(&vertex_list.d_dists, nodes);
        if (!vertex_list.d_dists_out) (void)mapgraph::device_alloc(&vertex_list.d_dists_out, nodes);
        (void)mapgraph::iota(vertex_list.d_dists, nodes);
        (void)mapgraph::iota(vertex_list.d_dists_out, nodes);
        (void)mapgraph::iota(d_frontier_keys[0], nodes);
        (void)mapgraph::iota(d_frontier_keys[1], nodes);
        (void)mapgraph::fill(d_frontier_values[0], (MiscType)INIT_VALUE, nodes);
        (void)mapgraph::fill(d_frontier_values[1], (MiscType)INIT_VALUE, nodes);
    }

    static void Release(VertexType &vertex_list, EdgeType &) {
        if (vertex_list.d_dists) (void)hipFree(vertex_list.d_dists);
        if (vertex_list.d_dists_out) (void)hipFree(vertex_list.d_dists_out);
        vertex_list.d_dists = vertex_list.d_dists_out = nullptr;
    }

    struct gather_vertex {
        __device__ void operator()(const int, const GatherType, VertexType &, EdgeType &) {}
    };

    struct gather_edge {
        __device__ void operator()(const int, const int, const int, VertexType &, EdgeType &, GatherType &) {}
    };

    struct gather_sum {
        __device__ GatherType operator()(const GatherType &, const GatherType &) { return 0; }
    };

    // cc.h:189-216
    struct apply {
        __device__ void operator()(const int vertex_id, const int iteration, GatherType gathervalue,
                                   VertexType &vertex_list, EdgeType &, char &changed) {
            const int oldvalue = vertex_list.d_dists[vertex_id];
            const int newvalue = min(oldvalue, gathervalue);
            changed = (iteration == 0 || oldvalue != newvalue) ? 1 : 0;
            vertex_list.d_dists_out[vertex_id] = newvalue;
        }
    };

    // cc.h:219-227
    struct post_apply {
        __device__ void operator()(const int vertex_id, VertexType &vertex_list, EdgeType &, GatherType *gather_tmp) {
            vertex_list.d_dists[vertex_id] = vertex_list.d_dists_out[vertex_id];
            gather_tmp[vertex_id] = INIT_VALUE;
        }
    };

    struct expand_vertex {
        __device__ bool operator()(const int, const char changed, VertexType &, EdgeType &) { return changed == 1; }
    };

    // cc.h:289-301: offer the own label to neighbours with a larger one
    struct expand_edge {
        __device__ void operator()(const bool changed, const int iteration, const int vertex_id,
                                   const int neighbor_id_in, const int, VertexType &vertex_list, EdgeType &,
                                   int &frontier, int &misc_value) {
            const int src_dist = vertex_list.d_dists[vertex_id];
            const int dst_dist = vertex_list.d_dists[neighbor_id_in];
            frontier = ((changed || iteration == 0) && dst_dist > src_dist) ? neighbor_id_in : -1;
            misc_value = src_dist;
        }
    };

    // cc.h:328-341
    struct contract {
        __device__ void operator()(const int, int &vertex_id, VertexType &, EdgeType &, GatherType *gather_tmp,
                                   int &misc_value) {
            atomicMin(&gather_tmp[vertex_id], misc_value);
        }
    };

    static void extractResult(VertexType &vertex_list, DataType *h_output) {
        (void)hipMemcpy(h_output, vertex_list.d_dists, sizeof(DataType) * vertex_list.nodes, hipMemcpyDeviceToHost);
    }
};
