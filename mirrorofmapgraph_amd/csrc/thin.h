// Thin searches (serving): one workgroup per query, the whole search in one launch.
#pragma once

#include "engine.h"

namespace mg {

constexpr int THIN_MAX_JOBS = 64;  // queries per launch (passed by value)

// Device state of a thin engine: per lane (a concurrent query) the result row, two
// frontier queues of n entries and, for SSSP, a far list and two bitmaps.
struct ThinArgs {
    const int *off, *adj, *w;
    int n, nwords;
    long long delta;    // SSSP bucket width
    int *rows;          // lanes x n: depths / distances, the result rows
    int *q;             // lanes x 2n
    int *far;           // lanes x n (SSSP)
    uint32_t *bits;     // lanes x 2 nwords (SSSP: queued-near, waiting-far)
    long long *stats;   // lanes x 2: traversed edges, reached vertices (pinned host memory)
};

struct ThinBatch {
    int n;
    int lane[THIN_MAX_JOBS];
    int src[THIN_MAX_JOBS];
};

struct ThinDev {
    int lanes = 0;
    int64_t bytes = 0;
    ThinArgs a{};
    long long *h_stats = nullptr;  // pinned host array the kernels write a.stats into
};

// Device bytes of a thin engine with `lanes` lanes.
int64_t thin_dev_bytes(const mg_graph *g, bool sssp, int lanes);
int thin_dev_alloc(ThinDev &t, mg_graph *g, bool sssp, int lanes, long long delta);
void thin_dev_free(ThinDev &t);
// Launches one workgroup per job on st.
int thin_launch(const ThinArgs &a, const ThinBatch &jobs, bool sssp, hipStream_t st);

}  // namespace mg
