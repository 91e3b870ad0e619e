// Thin searches (serving): one workgroup per query, the whole search in one launch.
#pragma once

#include "engine.h"

namespace mg {

constexpr int THIN_MAX_JOBS = 64;  // queries per launch (passed by value)
constexpr int THIN_STAT = 4;       // per-lane words of the host-visible status

// Device state of a thin engine: per lane (a concurrent query) the result row, two
// frontier queues of n entries and, for SSSP, a far list and two bitmaps, plus the
// few scalars of a paused search.
struct ThinArgs {
    const int *off, *adj, *w;
    int n, nwords;
    long long delta;    // SSSP bucket width
    int *rows;          // lanes x n: depths / distances, the result rows
    int *q;             // lanes x 2n
    int *far;           // lanes x n (SSSP)
    uint32_t *bits;     // lanes x 2 nwords (SSSP: queued-near, waiting-far)
    long long *stats;   // lanes x THIN_STAT in pinned host memory: traversed edges,
                        // reached vertices, finished (1) or paused (0)
    long long *state;   // lanes x 4: a paused search's queue parity, queue size, level
                        // (BFS) or far-list size (SSSP), bucket bound (SSSP)
};

// One launch: a workgroup per job.  A job starts a search from src, or (src = -1)
// resumes the lane's paused one; it pauses at a level boundary once it has run
// `slice` wall-clock ticks in this launch (0: runs to the end).  A finished search
// copies its row to out (a device address; null: none).
struct ThinBatch {
    int n;
    int lane[THIN_MAX_JOBS];
    int src[THIN_MAX_JOBS];
    void *out[THIN_MAX_JOBS];
    long long slice;
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
