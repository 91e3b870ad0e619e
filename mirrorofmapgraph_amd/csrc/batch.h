// Shared pieces of the batched searches: 256-bit per-vertex source masks, the
// bitmap walk used by every batched kernel, and the per-graph batch state.
#pragma once
#include "frontier.h"

namespace mg {

constexpr int SB_TASK = 32;  // batched SSSP: edges per wave task
constexpr int SB_RING = 16;  // batched SSSP: bucket bitmaps in the ring


struct M8 {
    uint4 a, b;
};

__device__ __forceinline__ M8 ld8(const uint4 *p, int v) { return M8{p[2 * v], p[2 * v + 1]}; }
__device__ __forceinline__ void st8(uint4 *p, int v, const M8 &m) {
    p[2 * v] = m.a;
    p[2 * v + 1] = m.b;
}
__device__ __forceinline__ M8 or8(const M8 &x, const M8 &y) {
    return M8{make_uint4(x.a.x | y.a.x, x.a.y | y.a.y, x.a.z | y.a.z, x.a.w | y.a.w),
              make_uint4(x.b.x | y.b.x, x.b.y | y.b.y, x.b.z | y.b.z, x.b.w | y.b.w)};
}
__device__ __forceinline__ M8 andnot8(const M8 &x, const M8 &y) {
    return M8{make_uint4(x.a.x & ~y.a.x, x.a.y & ~y.a.y, x.a.z & ~y.a.z, x.a.w & ~y.a.w),
              make_uint4(x.b.x & ~y.b.x, x.b.y & ~y.b.y, x.b.z & ~y.b.z, x.b.w & ~y.b.w)};
}
__device__ __forceinline__ bool any8(const M8 &x) {
    return (x.a.x | x.a.y | x.a.z | x.a.w | x.b.x | x.b.y | x.b.z | x.b.w) != 0;
}
__device__ __forceinline__ uint32_t word8(const M8 &x, int w) {
    switch (w) {
        case 0: return x.a.x;
        case 1: return x.a.y;
        case 2: return x.a.z;
        case 3: return x.a.w;
        case 4: return x.b.x;
        case 5: return x.b.y;
        case 6: return x.b.z;
        default: return x.b.w;
    }
}
__device__ __forceinline__ int popc8(const M8 &x) {
    return __popc(x.a.x) + __popc(x.a.y) + __popc(x.a.z) + __popc(x.a.w) + __popc(x.b.x) + __popc(x.b.y) +
           __popc(x.b.z) + __popc(x.b.w);
}
__device__ __forceinline__ M8 shfl_xor8(const M8 &x, int o) {
    return M8{make_uint4(__shfl_xor(x.a.x, o), __shfl_xor(x.a.y, o), __shfl_xor(x.a.z, o), __shfl_xor(x.a.w, o)),
              make_uint4(__shfl_xor(x.b.x, o), __shfl_xor(x.b.y, o), __shfl_xor(x.b.z, o), __shfl_xor(x.b.w, o))};
}

// Calls fn(v) for the set bits of a bitmap, block-uniformly: each round a
// workgroup compacts BLK/2 words into LDS and hands them out BLK at a time (fn is
// called by every thread, with v = -1 for padding).
template <class Fn>
__device__ void for_each_bit(const uint32_t *bm, int nwords, Fn fn) {
    constexpr int WPR = BLK / 2;
    __shared__ int s_q[WPR * 32];
    __shared__ int s_w[BLK / WAVE];
    for (int w0 = blockIdx.x * WPR; w0 < nwords; w0 += gridDim.x * WPR) {
        const int w = w0 + (int)threadIdx.x;
        uint32_t word = (threadIdx.x < WPR && w < nwords) ? bm[w] : 0u;
        int total;
        int k = block_exclusive_scan<BLK>(__popc(word), s_w, &total);
        while (word) {
            const int b = __ffs(word) - 1;
            s_q[k++] = w * 32 + b;
            word &= word - 1;
        }
        __syncthreads();
        for (int t0 = 0; t0 < total; t0 += BLK) {
            const int i = t0 + (int)threadIdx.x;
            fn(i < total ? s_q[i] : -1);
        }
        __syncthreads();
    }
}

int ensure_batch(mg_graph *g, bool sssp);
int check_batch_sources(const char *fn, const mg_graph *g, const int32_t *srcs, int32_t k);

}  // namespace mg
