// gas.h -- MapGraph GAS Program concept for the MI355X engine.
//
// A program is a struct that the enactor drives (same concept as the reference,
// doc/mapgraph-overview.h:107-251 and Algorithms/BFS/bfs.h:28-262):
//
//   typedef ... DataType, MiscType, GatherType;   typedef int VertexId, SizeT;
//   static const DataType INIT_VALUE;             static const bool allow_duplicates;
//   struct VertexType { ...device pointers...; int nodes, edges; };
//   struct EdgeType   { ...device pointers...; int nodes, edges; };
//   static void Initialize(int directed, int nodes, int edges, int num_srcs, int *srcs,
//       int *d_row_offsets, int *d_column_indices, int *d_column_offsets,
//       int *d_row_indices, DataType *d_edge_values, VertexType &, EdgeType &,
//       int *d_frontier_keys[3], MiscType *d_frontier_values[3]);
//   static SrcVertex srcVertex();  static GatherEdges gatherOverEdges();
//   static ApplyVertices applyOverEdges();  static PostApplyVertices postApplyOverEdges();
//   static ExpandEdges expandOverEdges();
//   __device__ functors gather_edge, gather_sum, apply, post_apply, expand_vertex,
//   expand_edge, contract (signatures as in the reference programs);
//   static void extractResult(VertexType &, DataType *h_output);
//
// The enumerations below keep the reference's names and order
// (GASengine/csr_problem.cuh:68-91) so programs written for MapGraph compile
// unchanged in their device half.  Their host half replaces cudaMalloc/cudaMemcpy/
// thrust with HIP and the helpers in namespace mapgraph.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstddef>
#include <cstdint>

enum SrcVertex { SINGLE, ALL };

enum GatherEdges { NO_GATHER_EDGES, GATHER_IN_EDGES, GATHER_OUT_EDGES, GATHER_ALL_EDGES };

enum ExpandEdges { NO_EXPAND_EDGES, EXPAND_IN_EDGES, EXPAND_OUT_EDGES, EXPAND_ALL_EDGES };

enum ApplyVertices { NO_APPLY_VERTICES, APPLY_ALL, APPLY_FRONTIER };

enum PostApplyVertices { NO_POST_APPLY_VERTICES, POST_APPLY_ALL, POST_APPLY_FRONTIER };

namespace mapgraph {

namespace detail {
template <class T>
__global__ void fill_kernel(T *p, T v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
template <class T>
__global__ void iota_kernel(T *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = (T)i;
}
// out[v] = off[v+1] - off[v] (+ off2[v+1] - off2[v] when off2 is given); a template
// so that every translation unit including this header can instantiate it
template <class T>
__global__ void degree_kernel(const T *off, const T *off2, T *out, int n) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        out[v] = off[v + 1] - off[v] + (off2 ? off2[v + 1] - off2[v] : 0);
}
inline unsigned grid_for(size_t n) {
    size_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}
}  // namespace detail

// Device allocation for program state (the reference's cudaMalloc).
template <class T>
inline hipError_t device_alloc(T **p, size_t count) {
    return hipMalloc((void **)p, (count ? count : 1) * sizeof(T));
}

// p[0..n) = v (the reference's b40c MemsetKernel).
template <class T>
inline hipError_t fill(T *p, T v, size_t n, hipStream_t s = 0) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(detail::fill_kernel<T>, dim3(detail::grid_for(n)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

// p[i] = i (the reference's b40c SequenceKernel).
template <class T>
inline hipError_t iota(T *p, size_t n, hipStream_t s = 0) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(detail::iota_kernel<T>, dim3(detail::grid_for(n)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

// Vertex degrees from offsets (the reference's thrust::adjacent_difference on row
// offsets, PageRank.h:147-161); with off2, the sum of both views' degrees.
inline hipError_t degrees(const int *d_off, const int *d_off2, int n, int *d_out, hipStream_t s = 0) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(detail::degree_kernel<int>, dim3(detail::grid_for(n)), dim3(256), 0, s, d_off, d_off2, d_out,
                       n);
    return hipGetLastError();
}

}  // namespace mapgraph
