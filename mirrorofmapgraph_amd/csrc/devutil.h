// Wave64 / workgroup building blocks for gfx950 (CDNA4) kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mg {

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Number of set bits of `mask` in lanes below the calling lane.
__device__ __forceinline__ int lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }

// Appends `item` for every lane with `want` to q[], reserving space with one atomic
// per wave on *counter.  Returns nothing; lanes without `want` write nothing.
template <typename T>
__device__ __forceinline__ void wave_append(bool want, T item, T *q, int *counter) {
    uint64_t mask = wave_ballot(want);
    if (!mask) return;
    int leader = __ffsll((unsigned long long)mask) - 1;
    int base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
    if (want) q[base + lanes_below(mask)] = item;
}

// Wave-wide sum (all active lanes get the result).
__device__ __forceinline__ long long wave_sum_ll(long long x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

__device__ __forceinline__ int wave_sum_i(int x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Exclusive scan across a workgroup of BLOCK threads (BLOCK a multiple of 64).
// `s_wave` must hold BLOCK/64 ints.  Returns the exclusive prefix; *total gets the
// workgroup sum.  Contains __syncthreads().
template <int BLOCK>
__device__ __forceinline__ int block_exclusive_scan(int x, int *s_wave, int *total) {
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    int incl = x;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == WAVE - 1) s_wave[wid] = incl;
    __syncthreads();
    int wave_off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / WAVE; ++w) {
        int s = s_wave[w];
        if (w < wid) wave_off += s;
        tot += s;
    }
    *total = tot;
    __syncthreads();
    return wave_off + incl - x;
}

// p[i] when ok, else dflt, with the load made unconditional (from p[safe] when !ok;
// safe must be in bounds).  A conditional load compiles to a branch that waits for
// the load before the next one; loads written this way issue back to back, so a
// batch of gathers is in flight together.
template <typename T, typename I>
__device__ __forceinline__ T load_or(const T *p, I i, bool ok, I safe, T dflt) {
    const T v = p[ok ? i : safe];
    return ok ? v : dflt;
}

// Inclusive minimum scan over runs of lanes: lane l gets the minimum of x over lanes
// [start, l], where start <= l is the first lane of l's run (runs are contiguous).
// DPP row shifts within rows of 16 lanes, then the row_bcast:15 / row_bcast:31 steps
// carry row ends across rows; lanes whose source lies outside the run ignore it.
__device__ __forceinline__ int seg_min_scan(int x, int lane, int start) {
    int y;
    y = __builtin_amdgcn_update_dpp(INT32_MAX, x, 0x111, 0xF, 0xF, false);  // row_shr:1
    if (lane - 1 >= start) x = y < x ? y : x;
    y = __builtin_amdgcn_update_dpp(INT32_MAX, x, 0x112, 0xF, 0xF, false);  // row_shr:2
    if (lane - 2 >= start) x = y < x ? y : x;
    y = __builtin_amdgcn_update_dpp(INT32_MAX, x, 0x114, 0xF, 0xF, false);  // row_shr:4
    if (lane - 4 >= start) x = y < x ? y : x;
    y = __builtin_amdgcn_update_dpp(INT32_MAX, x, 0x118, 0xF, 0xF, false);  // row_shr:8
    if (lane - 8 >= start) x = y < x ? y : x;
    y = __builtin_amdgcn_update_dpp(INT32_MAX, x, 0x142, 0xA, 0xF, false);  // row_bcast:15 -> rows 1, 3
    if ((lane & ~15) - 1 >= start && (lane & 16)) x = y < x ? y : x;
    y = __builtin_amdgcn_update_dpp(INT32_MAX, x, 0x143, 0xC, 0xF, false);  // row_bcast:31 -> rows 2, 3
    if (lane >= 32 && 31 >= start) x = y < x ? y : x;
    return x;
}

// Inclusive OR over lanes [start, lane] of this lane's run (every lane of the wave
// active; start <= lane).
__device__ __forceinline__ uint32_t seg_or_scan(uint32_t x, int lane, int start) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane - d >= start) x |= y;
    }
    return x;
}

__device__ __forceinline__ bool test_bit(const uint32_t *bm, int v) {
    return (bm[v >> 5] >> (v & 31)) & 1u;
}

// Atomically sets bit v; returns true when this call set it.
__device__ __forceinline__ bool set_bit_first(uint32_t *bm, int v) {
    const uint32_t bit = 1u << (v & 31);
    return !(atomicOr(&bm[v >> 5], bit) & bit);
}

}  // namespace mg
