// bfs -- breadth-first search as a GAS program (port of Algorithms/BFS/bfs.h:28-262).
// Label = iteration + 1 on first visit, -1 when unreached; duplicates allowed.
// Device functors keep the reference semantics; the host half uses HIP.
#pragma once

#include <vector>

#include "../gas.h"

struct bfs {
    typedef int DataType;
    typedef DataType MiscType;
    typedef DataType GatherType;
    typedef int VertexId;
    typedef int SizeT;

    static constexpr DataType INIT_VALUE = -1;
    static constexpr bool allow_duplicates = true;

    struct VertexType {
        int *d_labels = nullptr;
        int nodes = 0;
        int edges = 0;
    };

    struct EdgeType {
        int nodes = 0;
        int edges = 0;
    };

    // bfs.h:62-126: labels = -1, sources = 0, frontier = sources.  The label array
    // is allocated once and reused by later searches (the reference reallocates it
    // on every call, bfs.h:80).
    static void Initialize(const int directed, const int nodes, const int edges, int num_srcs, int *srcs,
                           int *d_row_offsets, int *d_column_indices, int *d_column_offsets, int *d_row_indices,
                           int *d_edge_values, VertexType &vertex_list, EdgeType &edge_list,
                           int *d_frontier_keys[3], MiscType *d_frontier_values[3]) {
        (void)directed, (void)d_row_offsets, (void)d_column_indices, (void)d_column_offsets, (void)d_row_indices,
            (void)d_edge_values;
        vertex_list.nodes = edge_list.nodes = nodes;
        vertex_list.edges = edge_list.edges = edges;
        if (!vertex_list.d_labels) (void)mapgraph::device_alloc(&vertex_list.d_labels, nodes);
        (void)mapgraph::fill(vertex_list.d_labels, -1, nodes);
        const int zero = 0;
        for (int i = 0; i < num_srcs; ++i)
            (void)hipMemcpy(vertex_list.d_labels + srcs[i], &zero, sizeof zero, hipMemcpyHostToDevice);
        (void)hipMemcpy(d_frontier_keys[0], srcs, num_srcs * sizeof(int), hipMemcpyHostToDevice);
        (void)hipMemcpy(d_frontier_keys[1], srcs, num_srcs * sizeof(int), hipMemcpyHostToDevice);
        std::vector<MiscType> zeros(num_srcs, 0);
        (void)hipMemcpy(d_frontier_values[0], zeros.data(), num_srcs * sizeof(MiscType), hipMemcpyHostToDevice);
        (void)hipMemcpy(d_frontier_values[1], zeros.data(), num_srcs * sizeof(MiscType), hipMemcpyHostToDevice);
    }

    static void Release(VertexType &vertex_list, EdgeType &) {
        if (vertex_list.d_labels) (void)hipFree(vertex_list.d_labels);
        vertex_list.d_labels = nullptr;
    }

    static SrcVertex srcVertex() { return SINGLE; }
    static GatherEdges gatherOverEdges() { return NO_GATHER_EDGES; }
    static ApplyVertices applyOverEdges() { return NO_APPLY_VERTICES; }
    static ExpandEdges expandOverEdges() { return EXPAND_OUT_EDGES; }
    static PostApplyVertices postApplyOverEdges() { return NO_POST_APPLY_VERTICES; }

    struct gather_sum {
        __device__ GatherType operator()(GatherType left, GatherType right) { return left + right; }
    };

    struct gather_vertex {
        __device__ void operator()(const int, const GatherType, VertexType &, EdgeType &) {}
    };

    struct gather_edge {
        __device__ void operator()(const int, const int, const int, VertexType &, EdgeType &, GatherType &) {}
    };

    struct apply {
        __device__ void operator()(const int, const int, GatherType, VertexType &, EdgeType &, char &changed) {
            changed = 1;
        }
    };

    struct post_apply {
        __device__ void operator()(const int, VertexType &, EdgeType &, GatherType *) {}
    };

    struct expand_vertex {
        __device__ bool operator()(const int, const char, VertexType &, EdgeType &) { return true; }
    };

    // bfs.h:216-227: every out-neighbour goes to the edge frontier
    struct expand_edge {
        __device__ void operator()(const bool, const int, const int, const int neighbor_id_in, const int,
                                   VertexType &, EdgeType &, int &frontier, int &) {
            frontier = neighbor_id_in;
        }
    };

    // bfs.h:229-255: first visit sets the label, later visits are dropped
    struct contract {
        __device__ void operator()(const int iteration, int &vertex_id, VertexType &vertex_list, EdgeType &,
                                   GatherType *, int &) {
            const int label = vertex_list.d_labels[vertex_id];
            if (label != -1) vertex_id = -1;
            else vertex_list.d_labels[vertex_id] = iteration + 1;
        }
    };

    static void extractResult(VertexType &vertex_list, DataType *h_output) {
        (void)hipMemcpy(h_output, vertex_list.d_labels, sizeof(DataType) * vertex_list.nodes, hipMemcpyDeviceToHost);
    }
};
