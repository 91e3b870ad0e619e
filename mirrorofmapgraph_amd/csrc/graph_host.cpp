// Host graph tools: seeded generators, CSR/CSC builders and the Matrix Market
// reader.  These feed the device graph store; nothing here touches the GPU.
//
// The CSR builder mirrors CsrGraph::FromCoo (b40c/graph/csr_graph.cuh:74-164) when no
// cleaning flag is given (stable grouping by row, input order kept inside a row) and
// adds the symmetrize / self-loop / dedupe / column-sort steps the R-MAT and grid
// workloads need.  The Matrix Market reader follows ReadMarketStream
// (b40c/graph/builder/market.cuh:42-166): 1-based ids, optional value column
// defaulting to 1, value cast to the integer value type, and no symmetrization.
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "mg_internal.h"

namespace mg {

static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

int host_threads() {
    const char *env = getenv("MG_HOST_THREADS");
    int t = env ? atoi(env) : (int)std::thread::hardware_concurrency();
    if (t <= 0) t = 1;
    return std::min(t, 16);
}

// Runs body(t, lo, hi) over [0, n) split into `nt` contiguous chunks.
template <class F>
static void parallel_chunks(int64_t n, int nt, F body) {
    if (nt <= 1 || n < (1 << 16)) {
        body(0, (int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        th.emplace_back([=] { body(t, lo, hi); });
    }
    for (auto &x : th) x.join();
}

// Runs body(t) for t in [0, nt) on nt threads.
template <class F>
static void parallel_tasks(int nt, F body) {
    if (nt <= 1) {
        body(0);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back([=] { body(t); });
    for (auto &x : th) x.join();
}

// Stable LSD radix sort of 64-bit keys on their low `bits` bits (11-bit digits),
// multi-threaded: per-thread histograms, global bucket-major prefix, scatter.
static void radix_sort_u64(std::vector<uint64_t> &keys, int bits) {
    const int64_t n = (int64_t)keys.size();
    if (n < 2) return;
    std::vector<uint64_t> tmp(n);
    const int R = 11, B = 1 << R;
    const int nt = host_threads();
    std::vector<int64_t> hist((size_t)nt * B);
    uint64_t *src = keys.data(), *dst = tmp.data();
    int passes = 0;
    for (int shift = 0; shift < bits; shift += R, ++passes) {
        std::fill(hist.begin(), hist.end(), 0);
        parallel_chunks(n, nt, [&](int t, int64_t lo, int64_t hi) {
            int64_t *h = &hist[(size_t)t * B];
            for (int64_t i = lo; i < hi; ++i) h[(src[i] >> shift) & (B - 1)]++;
        });
        int64_t run = 0;
        for (int d = 0; d < B; ++d)
            for (int t = 0; t < nt; ++t) {
                int64_t c = hist[(size_t)t * B + d];
                hist[(size_t)t * B + d] = run;
                run += c;
            }
        parallel_chunks(n, nt, [&](int t, int64_t lo, int64_t hi) {
            int64_t *h = &hist[(size_t)t * B];
            for (int64_t i = lo; i < hi; ++i) {
                uint64_t k = src[i];
                dst[h[(k >> shift) & (B - 1)]++] = k;
            }
        });
        std::swap(src, dst);
    }
    if (passes & 1) std::memcpy(keys.data(), src, sizeof(uint64_t) * n);
}

static int bits_for(int64_t n) {
    int b = 0;
    while (b < 62 && ((int64_t)1 << b) < n) ++b;
    return b;
}

}  // namespace mg

using namespace mg;

extern "C" {

size_t mg_last_error(char *buf, size_t len) {
    const std::string &e = g_last_error;
    if (buf && len) {
        size_t k = std::min(len - 1, e.size());
        std::memcpy(buf, e.data(), k);
        buf[k] = 0;
    }
    return e.size();
}

const char *mg_version(void) { return "0.1.0"; }

int64_t mg_rmat_num_edges(int32_t scale, int32_t edge_factor) {
    if (scale < 0 || scale > 30 || edge_factor < 0) return -1;
    return (int64_t)edge_factor << scale;
}

int mg_gen_rmat_edges(int32_t scale, int32_t edge_factor, double a, double b, double c,
                      uint64_t seed, int32_t *src, int32_t *dst) {
    clear_error();
    if (scale < 1 || scale > 30 || edge_factor < 1 || !src || !dst)
        return set_error(MG_ERR_INVALID_ARG, "mg_gen_rmat_edges: bad scale/edge_factor/pointer");
    if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0))
        return set_error(MG_ERR_INVALID_ARG, "mg_gen_rmat_edges: probabilities must be >= 0 and sum <= 1");
    // Integer thresholds on a 32-bit draw keep the quadrant choice exact and portable.
    const uint64_t ta = (uint64_t)(a * 4294967296.0 + 0.5);
    const uint64_t tab = (uint64_t)((a + b) * 4294967296.0 + 0.5);
    const uint64_t tabc = (uint64_t)((a + b + c) * 4294967296.0 + 0.5);
    const int64_t m = (int64_t)edge_factor << scale;
    const uint64_t ms = mix64(seed);
    parallel_chunks(m, host_threads(), [&](int, int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; ++k) {
            uint32_t u = 0, v = 0;
            for (int l = 0; l < scale; ++l) {
                uint64_t r = mix64(ms + ((uint64_t)k * 64 + (uint64_t)l) * 0x9e3779b97f4a7c15ULL) >> 32;
                uint32_t i = r >= tab, j = (r >= ta && r < tab) || r >= tabc;
                u = (u << 1) | i;
                v = (v << 1) | j;
            }
            src[k] = (int32_t)u;
            dst[k] = (int32_t)v;
        }
    });
    return MG_OK;
}

int64_t mg_grid_num_edges(int32_t width, int32_t height) {
    if (width < 1 || height < 1) return -1;
    return (int64_t)height * (width - 1) + (int64_t)(height - 1) * width;
}

int mg_gen_grid_edges(int32_t width, int32_t height, double delete_frac, uint64_t seed,
                      int32_t *src, int32_t *dst, int64_t *m_out) {
    clear_error();
    if (width < 1 || height < 1 || (int64_t)width * height > INT32_MAX || !src || !dst || !m_out ||
        !(delete_frac >= 0 && delete_frac <= 1))
        return set_error(MG_ERR_INVALID_ARG, "mg_gen_grid_edges: bad arguments");
    const int64_t E = mg_grid_num_edges(width, height);
    const int64_t H = (int64_t)height * (width - 1);  // horizontal edges come first
    const int64_t D = (int64_t)std::floor(delete_frac * (double)E + 1e-9);
    std::vector<uint64_t> key(E);
    const uint64_t ms = mix64(seed);
    for (int64_t e = 0; e < E; ++e) key[e] = mix64(ms + (uint64_t)e * 0x9e3779b97f4a7c15ULL);
    std::vector<char> gone(E, 0);
    if (D > 0) {
        // delete the D edges that come first in (key, e) order
        std::vector<int64_t> idx(E);
        std::iota(idx.begin(), idx.end(), 0);
        auto lt = [&](int64_t x, int64_t y) { return key[x] != key[y] ? key[x] < key[y] : x < y; };
        std::nth_element(idx.begin(), idx.begin() + (D - 1), idx.end(), lt);
        for (int64_t i = 0; i < D; ++i) gone[idx[i]] = 1;
    }
    int64_t k = 0;
    for (int64_t e = 0; e < E; ++e) {
        if (gone[e]) continue;
        int64_t u, v;
        if (e < H) {
            int64_t r = e / (width - 1), cc = e % (width - 1);
            u = r * width + cc;
            v = u + 1;
        } else {
            int64_t f = e - H;
            u = f;  // f = r * width + c with r < height - 1
            v = f + width;
        }
        src[k] = (int32_t)u;
        dst[k] = (int32_t)v;
        ++k;
    }
    *m_out = k;
    return MG_OK;
}

int mg_csr_from_edges(int32_t n, int64_t m_in, const int32_t *src, const int32_t *dst,
                      const int32_t *vals, uint32_t flags, int32_t *row_offsets,
                      int32_t *col_idx, int32_t *vals_out, int64_t capacity, int64_t *m_out) {
    clear_error();
    if (n < 0 || m_in < 0 || (m_in > 0 && (!src || !dst)) || !row_offsets || !m_out ||
        (m_in > 0 && !col_idx) || ((vals == nullptr) != (vals_out == nullptr)))
        return set_error(MG_ERR_INVALID_ARG, "mg_csr_from_edges: bad arguments");
    for (int64_t i = 0; i < m_in; ++i)
        if ((uint32_t)src[i] >= (uint32_t)n || (uint32_t)dst[i] >= (uint32_t)n)
            return set_error(MG_ERR_INVALID_ARG, "mg_csr_from_edges: edge %lld = (%d,%d) outside [0,%d)",
                             (long long)i, src[i], dst[i], n);
    const bool sym = flags & MG_CSR_SYMMETRIZE, loops = flags & MG_CSR_DROP_LOOPS,
               dedupe = flags & MG_CSR_DEDUPE, sortc = (flags & MG_CSR_SORT_COLS) || dedupe;
    const int64_t m2 = sym ? 2 * m_in : m_in;
    // Sort keys: (row, col) when columns are sorted, else (row, input position) so the
    // row grouping is stable and keeps input order inside a row.
    const int vb = bits_for(n), pb = bits_for(m2 > 0 ? m2 : 1);
    const int lowbits = sortc ? vb : pb;
    const int shift = sortc && vals ? vb + pb : lowbits;  // bits below the row id
    if (shift + vb > 63)
        return set_error(MG_ERR_INVALID_ARG, "mg_csr_from_edges: graph too large for 64-bit sort keys");
    std::vector<uint64_t> keys;
    keys.reserve(m2);
    std::vector<int32_t> cols_in, vals_in;  // per input position (only when !sortc or vals)
    if (!sortc || vals) {
        cols_in.resize(m2);
        if (vals) vals_in.resize(m2);
    }
    int64_t p = 0;
    for (int64_t i = 0; i < m_in; ++i) {
        for (int dir = 0; dir < (sym ? 2 : 1); ++dir) {
            uint32_t u = dir ? dst[i] : src[i], v = dir ? src[i] : dst[i];
            if (loops && u == v) continue;
            if (!cols_in.empty()) {
                cols_in[p] = (int32_t)v;
                if (vals) vals_in[p] = vals[i];
            }
            uint64_t low = sortc ? (uint64_t)v : (uint64_t)p;
            if (sortc && vals) low = ((uint64_t)v << pb) | (uint64_t)p;  // keep value order stable
            keys.push_back(((uint64_t)u << shift) | low);
            ++p;
        }
    }
    radix_sort_u64(keys, shift + vb);
    std::fill(row_offsets, row_offsets + n + 1, 0);
    int64_t k = 0;
    const uint64_t lowmask = (shift >= 64) ? ~0ULL : ((1ULL << shift) - 1);
    int64_t prev_u = -1, prev_v = -1;
    for (int64_t i = 0; i < (int64_t)keys.size(); ++i) {
        uint64_t key = keys[i];
        int64_t u = (int64_t)(key >> shift);
        uint64_t low = key & lowmask;
        int64_t v, pos = -1;
        if (sortc && vals) {
            v = (int64_t)(low >> pb);
            pos = (int64_t)(low & ((1ULL << pb) - 1));
        } else if (sortc) {
            v = (int64_t)low;
        } else {
            pos = (int64_t)low;
            v = cols_in[pos];
        }
        if (dedupe && u == prev_u && v == prev_v) continue;
        prev_u = u;
        prev_v = v;
        if (k >= capacity)
            return set_error(MG_ERR_INVALID_ARG, "mg_csr_from_edges: capacity %lld too small", (long long)capacity);
        if (k >= INT32_MAX)
            return set_error(MG_ERR_INVALID_ARG, "mg_csr_from_edges: more than 2^31-1 edges");
        col_idx[k] = (int32_t)v;
        if (vals) vals_out[k] = vals_in[pos];
        row_offsets[u + 1]++;
        ++k;
    }
    for (int32_t v = 0; v < n; ++v) row_offsets[v + 1] += row_offsets[v];
    *m_out = k;
    return MG_OK;
}

int mg_edge_weights_hash(int32_t n, const int32_t *row_offsets, const int32_t *col_idx,
                         int64_t m, uint64_t seed, int32_t wmax, int32_t *w_out) {
    clear_error();
    if (n < 0 || !row_offsets || (m > 0 && (!col_idx || !w_out)) || wmax < 1 || row_offsets[n] != m)
        return set_error(MG_ERR_INVALID_ARG, "mg_edge_weights_hash: bad arguments");
    const uint64_t ms = mix64(seed);
    parallel_chunks(n, host_threads(), [&](int, int64_t lo, int64_t hi) {
        for (int64_t u = lo; u < hi; ++u)
            for (int64_t e = row_offsets[u]; e < row_offsets[u + 1]; ++e) {
                uint64_t v = (uint32_t)col_idx[e];
                uint64_t a = std::min<uint64_t>(u, v), b = std::max<uint64_t>(u, v);
                w_out[e] = 1 + (int32_t)(mix64(ms + ((a << 32) | b) * 0x9e3779b97f4a7c15ULL) % (uint64_t)wmax);
            }
    });
    return MG_OK;
}

int mg_csc_from_csr(int32_t n, int64_t m, const int32_t *row_offsets, const int32_t *col_idx,
                    int32_t *col_offsets, int32_t *row_idx, int32_t *csr_edge_id) {
    clear_error();
    if (n < 0 || !row_offsets || !col_offsets || (m > 0 && (!col_idx || !row_idx)) || row_offsets[n] != m)
        return set_error(MG_ERR_INVALID_ARG, "mg_csc_from_csr: bad arguments");
    std::fill(col_offsets, col_offsets + n + 1, 0);
    for (int64_t e = 0; e < m; ++e) col_offsets[col_idx[e] + 1]++;
    for (int32_t v = 0; v < n; ++v) col_offsets[v + 1] += col_offsets[v];
    std::vector<int64_t> fill(col_offsets, col_offsets + n);
    for (int32_t u = 0; u < n; ++u)
        for (int64_t e = row_offsets[u]; e < row_offsets[u + 1]; ++e) {
            int64_t slot = fill[col_idx[e]]++;
            row_idx[slot] = u;
            if (csr_edge_id) csr_edge_id[slot] = (int32_t)e;
        }
    return MG_OK;
}

// Degree order: vertex ids sorted by out-degree, highest first, ties by id (a stable
// counting sort over the degrees, or a stable comparison sort when degrees exceed 4n).
// Then two counting-sort transpositions: the first enumerates the renamed rows in
// order and files each edge under its renamed target (so every target's sources come
// out ascending), the second turns that back into rows whose targets are ascending.
int mg_csr_relabel_degree(int32_t n, int64_t m, const int32_t *row_offsets, const int32_t *col_idx,
                          const int32_t *vals, int32_t *new_of_old, int32_t *row_offsets_out,
                          int32_t *col_idx_out, int32_t *vals_out) {
    clear_error();
    if (n < 0 || m < 0 || !row_offsets || !new_of_old || !row_offsets_out || row_offsets[0] != 0 ||
        row_offsets[n] != m || (m > 0 && (!col_idx || !col_idx_out)) || (vals && m > 0 && !vals_out))
        return set_error(MG_ERR_INVALID_ARG, "mg_csr_relabel_degree: bad arguments");
    std::vector<int32_t> order((size_t)n);
    int64_t maxdeg = 0;
    for (int32_t v = 0; v < n; ++v) maxdeg = std::max<int64_t>(maxdeg, row_offsets[v + 1] - row_offsets[v]);
    if (maxdeg <= 4 * (int64_t)n + 16) {
        std::vector<int64_t> start((size_t)maxdeg + 2, 0);  // by degree, descending
        for (int32_t v = 0; v < n; ++v) start[(size_t)(maxdeg - (row_offsets[v + 1] - row_offsets[v])) + 1]++;
        for (int64_t d = 0; d <= maxdeg; ++d) start[(size_t)d + 1] += start[(size_t)d];
        for (int32_t v = 0; v < n; ++v) order[(size_t)start[(size_t)(maxdeg - (row_offsets[v + 1] - row_offsets[v]))]++] = v;
    } else {
        for (int32_t v = 0; v < n; ++v) order[(size_t)v] = v;
        std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
            return row_offsets[x + 1] - row_offsets[x] > row_offsets[y + 1] - row_offsets[y];
        });
    }
    for (int32_t i = 0; i < n; ++i) new_of_old[order[(size_t)i]] = i;
    // renamed rows i -> files of renamed targets j (sources ascending within a file).
    // Thread t owns the files of a contiguous range of targets (balanced by edges) and
    // scans every edge, so each file is filled in row order without synchronisation.
    std::vector<int32_t> ncol((size_t)m);
    parallel_chunks(m, host_threads(), [&](int, int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; ++e) ncol[(size_t)e] = new_of_old[col_idx[e]];
    });
    std::vector<int64_t> tstart((size_t)n + 1, 0);
    for (int64_t e = 0; e < m; ++e) tstart[(size_t)ncol[(size_t)e] + 1]++;
    for (int32_t j = 0; j < n; ++j) tstart[(size_t)j + 1] += tstart[(size_t)j];
    std::vector<int32_t> tsrc((size_t)m), tval(vals ? (size_t)m : 0);
    const int nt = m < (1 << 20) ? 1 : host_threads();
    // first id of each thread's range: ranges of about m / nt entries along `starts`
    auto split = [&](const int64_t *starts) {
        std::vector<int32_t> cut((size_t)nt + 1, n);
        cut[0] = 0;
        for (int t = 1; t < nt; ++t)
            cut[(size_t)t] = (int32_t)(std::upper_bound(starts, starts + n + 1, m * t / nt) - starts - 1);
        for (int t = 1; t <= nt; ++t) cut[(size_t)t] = std::max(cut[(size_t)t], cut[(size_t)t - 1]);
        return cut;
    };
    {
        const std::vector<int32_t> cut = split(tstart.data());
        parallel_tasks(nt, [&](int t) {
            {
                const int32_t jlo = cut[(size_t)t], jhi = cut[(size_t)t + 1];
                std::vector<int64_t> fill(tstart.begin() + jlo, tstart.begin() + jhi);
                for (int32_t i = 0; i < n; ++i) {
                    const int32_t v = order[(size_t)i];
                    for (int64_t e = row_offsets[v]; e < row_offsets[v + 1]; ++e) {
                        const int32_t j = ncol[(size_t)e];
                        if (j < jlo || j >= jhi) continue;
                        const int64_t slot = fill[(size_t)(j - jlo)]++;
                        tsrc[(size_t)slot] = i;
                        if (vals) tval[(size_t)slot] = vals[e];
                    }
                }
            }
        });
    }
    // back to rows: row i holds its targets j ascending (thread t owns a range of rows)
    row_offsets_out[0] = 0;
    for (int32_t i = 0; i < n; ++i) {
        const int32_t v = order[(size_t)i];
        row_offsets_out[i + 1] = row_offsets_out[i] + (row_offsets[v + 1] - row_offsets[v]);
    }
    {
        std::vector<int64_t> rstart(row_offsets_out, row_offsets_out + n + 1);
        const std::vector<int32_t> cut = split(rstart.data());
        parallel_tasks(nt, [&](int t) {
            {
                const int32_t ilo = cut[(size_t)t], ihi = cut[(size_t)t + 1];
                std::vector<int64_t> fill(rstart.begin() + ilo, rstart.begin() + ihi);
                for (int32_t j = 0; j < n; ++j)
                    for (int64_t k = tstart[(size_t)j]; k < tstart[(size_t)j + 1]; ++k) {
                        const int32_t i = tsrc[(size_t)k];
                        if (i < ilo || i >= ihi) continue;
                        const int64_t slot = fill[(size_t)(i - ilo)]++;
                        col_idx_out[slot] = j;
                        if (vals) vals_out[slot] = tval[(size_t)k];
                    }
            }
        });
    }
    return MG_OK;
}

int mg_pick_sources(int32_t n, const int32_t *row_offsets, int32_t k, uint64_t seed, int32_t *out) {
    clear_error();
    if (n <= 0 || !row_offsets || k < 0 || (k > 0 && !out))
        return set_error(MG_ERR_INVALID_ARG, "mg_pick_sources: bad arguments");
    bool any = false;
    for (int32_t v = 0; v < n && !any; ++v) any = row_offsets[v + 1] > row_offsets[v];
    if (!any && k > 0) return set_error(MG_ERR_INVALID_ARG, "mg_pick_sources: graph has no edges");
    uint64_t j = 0;
    for (int32_t i = 0; i < k;) {
        int32_t v = (int32_t)(hash_ctr(seed, j++) % (uint64_t)n);
        if (row_offsets[v + 1] > row_offsets[v]) out[i++] = v;
    }
    return MG_OK;
}

static int mtx_open(const char *path, FILE **f, int32_t *n, int64_t *nnz, int32_t *symmetric) {
    *f = fopen(path, "r");
    if (!*f) return set_error(MG_ERR_IO, "cannot open '%s': %s", path, strerror(errno));
    char line[1024];
    bool first = true;
    if (symmetric) *symmetric = 0;
    while (fgets(line, sizeof line, *f)) {
        if (first && !strncmp(line, "%%MatrixMarket", 14)) {
            if (symmetric && strstr(line, "symmetric")) *symmetric = 1;
        }
        first = false;
        if (line[0] == '%' || line[0] == '\n' || line[0] == '\r') continue;
        long long x, y, z;
        if (sscanf(line, "%lld %lld %lld", &x, &y, &z) != 3) {
            fclose(*f);
            return set_error(MG_ERR_FORMAT, "%s: invalid problem description", path);
        }
        if (x != y) {
            fclose(*f);
            return set_error(MG_ERR_FORMAT, "%s: not square (%lld, %lld)", path, x, y);
        }
        if (x < 0 || x > INT32_MAX || z < 0) {
            fclose(*f);
            return set_error(MG_ERR_FORMAT, "%s: sizes out of range", path);
        }
        *n = (int32_t)x;
        *nnz = z;
        return MG_OK;
    }
    fclose(*f);
    return set_error(MG_ERR_FORMAT, "%s: no graph found", path);
}

int mg_mtx_read_header(const char *path, int32_t *n, int64_t *nnz, int32_t *symmetric) {
    clear_error();
    if (!path || !n || !nnz) return set_error(MG_ERR_INVALID_ARG, "mg_mtx_read_header: null argument");
    FILE *f;
    int rc = mtx_open(path, &f, n, nnz, symmetric);
    if (rc) return rc;
    fclose(f);
    return MG_OK;
}

int mg_mtx_read(const char *path, int32_t *src, int32_t *dst, int32_t *val, int64_t capacity,
                int64_t *nnz_out) {
    clear_error();
    if (!path || !src || !dst || !nnz_out) return set_error(MG_ERR_INVALID_ARG, "mg_mtx_read: null argument");
    FILE *f;
    int32_t n;
    int64_t nnz;
    int rc = mtx_open(path, &f, &n, &nnz, nullptr);
    if (rc) return rc;
    if (nnz > capacity) {
        fclose(f);
        return set_error(MG_ERR_INVALID_ARG, "mg_mtx_read: capacity %lld < %lld entries", (long long)capacity,
                         (long long)nnz);
    }
    char line[1024];
    int64_t k = 0;
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '%' || line[0] == '\n' || line[0] == '\r') continue;
        if (k >= nnz) {
            fclose(f);
            return set_error(MG_ERR_FORMAT, "%s: encountered more than %lld edges", path, (long long)nnz);
        }
        long long r, c;
        double v = 1;
        int got = sscanf(line, "%lld %lld %lf", &r, &c, &v);
        if (got != 2 && got != 3) {
            fclose(f);
            return set_error(MG_ERR_FORMAT, "%s: badly formed edge %lld", path, (long long)k);
        }
        if (r < 1 || r > n || c < 1 || c > n) {
            fclose(f);
            return set_error(MG_ERR_FORMAT, "%s: edge %lld (%lld,%lld) outside [1,%d]", path, (long long)k, r, c, n);
        }
        src[k] = (int32_t)(r - 1);
        dst[k] = (int32_t)(c - 1);
        if (val) val[k] = (int32_t)v;
        ++k;
    }
    fclose(f);
    if (k != nnz) return set_error(MG_ERR_FORMAT, "%s: only %lld/%lld edges read", path, (long long)k, (long long)nnz);
    *nnz_out = k;
    return MG_OK;
}

}  // extern "C"
