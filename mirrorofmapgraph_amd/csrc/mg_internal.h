// Internal helpers shared by the host and device halves of libmapgraph_hip.so.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/mapgraph_c.h"

namespace mg {

// Records the last error of the calling thread; returns `code` so call sites can
// write `return set_error(MG_ERR_INVALID_ARG, "...")`.
int set_error(int code, const char *fmt, ...);
void clear_error();

// splitmix64 finalizer (Steele, Lea, Flood 2014): a bijective 64-bit mixer.
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Counter-based draw for stream position `ctr` under `seed`.
static inline uint64_t hash_ctr(uint64_t seed, uint64_t ctr) {
    return mix64(mix64(seed) + ctr * 0x9e3779b97f4a7c15ULL);
}

int host_threads();

}  // namespace mg
