"""GPU parity: every search runs on the device through the C-ABI and is compared with
the CPU oracle on the same seeded inputs.  BFS depths, SSSP distances and CC labels
must be bit-exact; PageRank is compared in normalised units p = r / n with a
max-abs tolerance of 1e-6 (the FP32 bar), and its error against the float64
oracle is checked too."""
import numpy as np
import pytest

from conftest import load_mtx, random_graph

pytestmark = pytest.mark.gpu
INT_MAX = 2**31 - 1
PR_TOL = 1e-6  # max-abs in normalised units


def _graph(mg, g, **kw):
    return mg.Graph(g, **kw)


def _check_bfs(mg, oracle, g, src, directed=1, **gkw):
    with _graph(mg, g, **gkw) as dg:
        got, st = dg.bfs(src, mg.options(directed=directed), stats=True)
    csc = g.csc() if directed == 0 else None
    exp = oracle.bfs(g.row_offsets, g.col_idx, src, directed=directed, csc=csc)
    assert np.array_equal(got, exp), f"BFS mismatch at {np.nonzero(got != exp)[0][:10]}"
    assert st.edges_traversed == oracle.teps_edges(g.row_offsets, exp)
    assert st.vertices_reached == int((exp >= 0).sum())
    return st


# ---------------------------------------------------------------- fixtures
@pytest.mark.parametrize("name", ["small.mtx", "trivialGraph.mtx", "chesapeake.mtx"])
def test_fixtures_bfs_sssp_cc(mg, oracle, device, name):
    g = load_mtx(mg, name)
    for directed in (1, 0):
        with mg.Graph(g, build_csc=True) as dg:
            for src in range(min(g.n, 5)):
                d = dg.bfs(src, mg.options(directed=directed))
                assert np.array_equal(d, oracle.bfs(g.row_offsets, g.col_idx, src, directed=directed))
                s = dg.sssp(src, mg.options(directed=directed))
                assert np.array_equal(s, oracle.sssp(g.row_offsets, g.col_idx, g.values, src, directed=directed))
            lab = dg.cc()
            assert np.array_equal(lab, oracle.cc(g.row_offsets, g.col_idx))


def test_small_known_answers_on_gpu(mg, device):
    g = load_mtx(mg, "small.mtx")
    with mg.Graph(g) as dg:
        assert dg.bfs(0).tolist() == [0, 1, 1, 1, 2, 2, 2]
        assert dg.sssp(0).tolist() == [0, 1, 4, 1, 2, 2, 2]
        assert dg.cc().tolist() == [0] * 7


# ---------------------------------------------------------------- random graphs
@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("symmetric", [True, False])
def test_bfs_random(mg, oracle, device, seed, symmetric):
    g = random_graph(mg, 5000, 40000, seed, symmetric=symmetric)
    src = int(np.argmax(g.degrees()))
    _check_bfs(mg, oracle, g, src, symmetric=symmetric)
    if not symmetric:  # directed search with the CSC view enables bottom-up
        _check_bfs(mg, oracle, g, src, build_csc=True)
        _check_bfs(mg, oracle, g, src, directed=0, build_csc=True)


@pytest.mark.parametrize("seed", [3, 4])
@pytest.mark.parametrize("symmetric", [True, False])
def test_sssp_random(mg, oracle, device, seed, symmetric):
    g = random_graph(mg, 4000, 30000, seed, symmetric=symmetric, wmax=64)
    src = int(np.argmax(g.degrees()))
    with mg.Graph(g, build_csc=True) as dg:
        got = dg.sssp(src)
        assert np.array_equal(got, oracle.sssp(g.row_offsets, g.col_idx, g.values, src))
        got0 = dg.sssp(src, mg.options(directed=0))
        assert np.array_equal(got0, oracle.sssp(g.row_offsets, g.col_idx, g.values, src, directed=0))


@pytest.mark.parametrize("symmetric", [True, False])
def test_cc_random(mg, oracle, device, symmetric):
    g = random_graph(mg, 20000, 15000, 5, symmetric=symmetric)
    with mg.Graph(g, symmetric=symmetric) as dg:
        assert np.array_equal(dg.cc(), oracle.cc(g.row_offsets, g.col_idx))


def _structured_cc_graph(mg, dense):
    """Symmetric graphs whose first neighbours are often larger than the vertex itself.  dense: three
    bipartite blocks (every vertex of a block's lower half joined to 12 of its upper half, so each lower
    vertex's neighbours are all larger) and isolated vertices, mean degree >= 8; sparse: a star centred on
    the last id, two paths, a descending chain and isolated vertices."""
    rng = np.random.default_rng(5)
    src, dst = [], []
    if dense:
        n, base = 3 * 1200 + 50, 0
        for _ in range(3):
            lo = np.arange(base, base + 600)
            for _ in range(12):
                src.append(lo)
                dst.append(rng.integers(base + 600, base + 1200, lo.size))
            base += 1200
    else:
        n = 5000
        src.append(np.arange(0, 1000))
        dst.append(np.full(1000, n - 1))
        src.append(np.arange(1000, 1999))
        dst.append(np.arange(1001, 2000))
        src.append(np.arange(3999, 3000, -1))
        dst.append(np.arange(3998, 2999, -1))
        src.append(np.arange(2000, 2500))
        dst.append(np.arange(2001, 2501)[::-1])
    s_, d_ = np.concatenate(src).astype(np.int32), np.concatenate(dst).astype(np.int32)
    flags = mg.MG_CSR_DEDUPE | mg.MG_CSR_SORT_COLS | mg.MG_CSR_DROP_LOOPS | mg.MG_CSR_SYMMETRIZE
    return mg.csr_from_edges(n, s_, d_, flags)


@pytest.mark.parametrize("dense", [True, False])
def test_cc_structured(mg, oracle, device, dense):
    g = _structured_cc_graph(mg, dense)
    assert (g.m >= 8 * g.n) == dense  # the plain-store first pass, or union-find rounds
    with mg.Graph(g, symmetric=True) as dg:
        assert np.array_equal(dg.cc(), oracle.cc(g.row_offsets, g.col_idx))


# ---------------------------------------------------------------- R-MAT / grid
@pytest.fixture(scope="module")
def s16(mg):
    return mg.rmat_graph(16, 16, seed=42, weight_max=64)


def test_rmat16_bfs_from_0(mg, oracle, device, s16):
    st = _check_bfs(mg, oracle, s16, 0, symmetric=True)
    assert st.bottom_up_levels > 0  # the direction switch is exercised
    # top-down only (no in-edge view) gives the same depths
    _check_bfs(mg, oracle, s16, 0)


def test_rmat16_bfs_many_sources(mg, oracle, device, s16):
    srcs = mg.pick_sources(s16, 8, 7).tolist() + [int(np.argmin(s16.degrees()))]  # incl. an isolated vertex
    with mg.Graph(s16, symmetric=True) as dg:
        for s in srcs:
            assert np.array_equal(dg.bfs(s), oracle.bfs(s16.row_offsets, s16.col_idx, s))


def test_rmat16_sssp_cc(mg, oracle, device, s16):
    with mg.Graph(s16, symmetric=True) as dg:
        d, st = dg.sssp(0, stats=True)
        exp = oracle.sssp(s16.row_offsets, s16.col_idx, s16.values, 0)
        assert np.array_equal(d, exp)
        assert st.edges_traversed == oracle.teps_edges(s16.row_offsets, exp, INT_MAX)
        assert np.array_equal(dg.cc(), oracle.cc(s16.row_offsets, s16.col_idx))


def test_grid_bfs_sssp(mg, oracle, device):
    g = mg.grid_graph(256, seed=3, weight_max=16)
    with mg.Graph(g, symmetric=True) as dg:
        d, st = dg.bfs(0, stats=True)
        exp = oracle.bfs(g.row_offsets, g.col_idx, 0)
        assert np.array_equal(d, exp)
        assert st.iterations >= exp.max()
        s = dg.sssp(0)
        assert np.array_equal(s, oracle.sssp(g.row_offsets, g.col_idx, g.values, 0))
        assert np.array_equal(dg.cc(), oracle.cc(g.row_offsets, g.col_idx))


@pytest.mark.parametrize("delta", [0, 1, 5, 64, -1])
def test_sssp_bucket_widths(mg, oracle, device, delta):
    """Every bucket width of the narrow-frontier loop (0 = auto, 1 = Dijkstra-like
    single-distance buckets, -1 = unbucketed Bellman-Ford) gives the oracle's
    distances, on a grid (all levels narrow) and on R-MAT (narrow and whole-GPU levels
    mixed, so the far list is handed over)."""
    grid = mg.grid_graph(128, seed=9, weight_max=16)
    rm = mg.rmat_graph(14, 8, seed=11, weight_max=64)
    for g, src in ((grid, 0), (grid, 128 * 64 + 64), (rm, 0), (rm, int(mg.pick_sources(rm, 1, 3)[0]))):
        with mg.Graph(g, symmetric=True) as dg:
            got, st = dg.sssp(src, mg.options(sssp_delta=delta), stats=True)
        exp = oracle.sssp(g.row_offsets, g.col_idx, g.values, src)
        assert np.array_equal(got, exp), (delta, src, np.nonzero(got != exp)[0][:10])
        assert st.edges_traversed == oracle.teps_edges(g.row_offsets, exp, INT_MAX)


@pytest.mark.parametrize("wmax", [1, 2, 3])
@pytest.mark.parametrize("relabel", [False, True])
def test_sssp_zero_and_tiny_weights(mg, oracle, device, wmax, relabel):
    """Weights drawn from [0, wmax) (zero-weight edges included) on R-MAT scale 15, where
    the dense pull levels run: the distance bound of the pull (fmin + w < dist) and the
    skip of vertices at or below fmin + (smallest weight) must keep every improvement."""
    g = mg.rmat_graph(15, 16, seed=21, weight_max=wmax)
    g.values = (g.values - 1).astype(np.int32) if wmax > 1 else np.zeros(g.m, np.int32)
    # keep the weights symmetric: hashed weights already are, the shift keeps that
    for src in (0, int(mg.pick_sources(g, 1, 5)[0])):
        with mg.Graph(g, symmetric=True, relabel=relabel) as dg:
            got, st = dg.sssp(src, stats=True)
        exp = oracle.sssp(g.row_offsets, g.col_idx, g.values, src)
        assert np.array_equal(got, exp), (wmax, src, np.nonzero(got != exp)[0][:10])
        assert st.edges_traversed == oracle.teps_edges(g.row_offsets, exp, INT_MAX)


# ---------------------------------------------------------------- limits and errors
def test_iter_num_limits_levels(mg, oracle, device, s16):
    with mg.Graph(s16, symmetric=True) as dg:
        full = oracle.bfs(s16.row_offsets, s16.col_idx, 0)
        for k in (1, 2):
            got = dg.bfs(0, mg.options(iter_num=k))
            exp = np.where(full <= k, full, -1)
            assert np.array_equal(got, exp)
        assert dg.bfs(0, mg.options(iter_num=0)).tolist() == [0] + [-1] * (s16.n - 1)


def test_tiny_and_empty_graphs(mg, device):
    g1 = mg.HostCSR(np.array([0, 0], np.int32), np.array([], np.int32), np.array([], np.int32))
    with mg.Graph(g1, weighted=True) as dg:
        assert dg.bfs(0).tolist() == [0]
        assert dg.sssp(0).tolist() == [0]
        assert dg.cc().tolist() == [0]
    g2 = mg.HostCSR(np.zeros(6, np.int32), np.array([], np.int32))
    with mg.Graph(g2, symmetric=True) as dg:
        assert dg.bfs(3).tolist() == [-1, -1, -1, 0, -1]
        assert dg.cc().tolist() == [0, 1, 2, 3, 4]


def test_errors_are_status_codes(mg, device, s16):
    unweighted = mg.HostCSR(s16.row_offsets, s16.col_idx)
    with mg.Graph(unweighted) as dg:
        with pytest.raises(mg.MgError) as e:
            dg.bfs(s16.n)
        assert e.value.code == mg.MG_ERR_INVALID_ARG
        with pytest.raises(mg.MgError) as e:
            dg.sssp(0)
        assert e.value.code == mg.MG_ERR_UNSUPPORTED
        with pytest.raises(mg.MgError) as e:
            dg.bfs(0, mg.options(directed=0))
        assert e.value.code == mg.MG_ERR_UNSUPPORTED
        with pytest.raises(mg.MgError) as e:
            dg.pagerank()
        assert e.value.code == mg.MG_ERR_UNSUPPORTED
    neg = mg.HostCSR(np.array([0, 1, 1], np.int32), np.array([1], np.int32), np.array([-3], np.int32))
    with mg.Graph(neg) as dg:
        with pytest.raises(mg.MgError) as e:
            dg.sssp(0)
        assert e.value.code == mg.MG_ERR_UNSUPPORTED


def test_repeated_searches_reuse_state(mg, oracle, device, s16):
    with mg.Graph(s16, symmetric=True) as dg:
        b0 = dg.device_bytes()
        exp = oracle.bfs(s16.row_offsets, s16.col_idx, 5)
        for _ in range(5):
            assert np.array_equal(dg.bfs(5), exp)
            dg.sssp(5)
            dg.cc()
        assert dg.device_bytes() == b0


def test_searches_of_different_depths_in_a_row(mg, oracle, device, s16):
    """A search queues as many steps as the previous one with the same options ran
    before it reads the control block; deeper, shallower and empty searches in a row
    must still run to the end."""
    g = mg.grid_graph(128, seed=5, weight_max=16)
    iso = int(np.argmin(s16.degrees()))
    for h, srcs in ((g, [0, 64 * 128 + 64, 0, 128 * 128 - 1, 5]), (s16, [0, iso, 0, 777, iso, 3])):
        with mg.Graph(h, symmetric=True) as dg:
            for s in srcs:
                assert np.array_equal(dg.bfs(s), oracle.bfs(h.row_offsets, h.col_idx, s)), s
                assert np.array_equal(dg.sssp(s), oracle.sssp(h.row_offsets, h.col_idx, h.values, s)), s


def test_device_csc_matches_host(mg, device):
    g = mg.rmat_graph(12, seed=2, symmetric=False)
    with mg.Graph(g, build_csc=True) as dg:
        co, ri, ce = dg.csc()
    hco, hri, hce = g.csc()
    assert np.array_equal(co, hco) and np.array_equal(ri, hri) and np.array_equal(ce, hce)


# ---------------------------------------------------------------- PageRank
def _pr_err(got, exp, n):
    return float(np.max(np.abs(got.astype(np.float64) - exp.astype(np.float64)))) / n


def test_pagerank_all_vertex_jacobi(mg, oracle, device):
    g = mg.rmat_graph(14, 16, seed=42, symmetric=False)
    with mg.Graph(g, build_csc=True) as dg:
        got, st = dg.pagerank(mg.options(pr_mode=mg.MG_PR_ALL, iter_num=50), stats=True)
    assert st.iterations == 50
    exp, it, _ = oracle.pr_jacobi_f32(g.row_offsets, g.col_idx, iters=50)
    assert _pr_err(got, exp, g.n) <= PR_TOL
    ref64, _, _ = oracle.pr_jacobi_f64(g.row_offsets, g.col_idx, iters=50)
    assert _pr_err(got, ref64, g.n) <= PR_TOL
    # low in-degree vertices are summed in the oracle's order: bit-exact there
    indeg = np.bincount(g.col_idx, minlength=g.n)
    small = indeg <= 32
    assert np.max(np.abs(got[small] - exp[small]) / np.maximum(exp[small], 1e-30)) < 1e-5


def test_pagerank_gas_program(mg, oracle, device):
    g = mg.rmat_graph(14, 16, seed=7, symmetric=False)
    with mg.Graph(g, build_csc=True) as dg:
        got, st = dg.pagerank(stats=True)
    exp, it = oracle.pr_gas(g.row_offsets, g.col_idx)
    assert _pr_err(got, exp, g.n) <= PR_TOL
    assert abs(st.iterations - it) <= 1


def test_pagerank_tolerance_stop(mg, oracle, device):
    g = mg.rmat_graph(12, 16, seed=5, symmetric=False)
    with mg.Graph(g, build_csc=True) as dg:
        got, st = dg.pagerank(mg.options(pr_mode=mg.MG_PR_ALL, iter_num=1000, pr_tol=1e-6), stats=True)
    exp, it, l1 = oracle.pr_jacobi_f32(g.row_offsets, g.col_idx, iters=1000, tol=1e-6)
    assert it < 1000 and abs(st.iterations - it) <= 1
    assert _pr_err(got, exp, g.n) <= PR_TOL


def test_pagerank_symmetric_graph_uses_csr(mg, oracle, device):
    g = mg.rmat_graph(12, 8, seed=3)
    with mg.Graph(g, symmetric=True) as dg:
        got = dg.pagerank(mg.options(pr_mode=mg.MG_PR_ALL, iter_num=20))
    exp, _, _ = oracle.pr_jacobi_f32(g.row_offsets, g.col_idx, iters=20)
    assert _pr_err(got, exp, g.n) <= PR_TOL


# ---------------------------------------------------------------- BFS predecessors
def _pred_reference(g, depth, csc=None):
    """predextract.cuh:131-139 + ReduceFunctor max: over in-edges (CSC), the largest
    neighbour u with depth[v] - depth[u] == 1, else -1."""
    co, ri, _ = csc if csc is not None else g.csc()
    n = g.n
    dst = np.repeat(np.arange(n), np.diff(co))
    ok = (depth[dst].astype(np.int64) - depth[ri].astype(np.int64)) == 1
    pred = np.full(n, -1, np.int64)
    np.maximum.at(pred, dst[ok], ri[ok].astype(np.int64))
    return pred.astype(np.int32)


@pytest.mark.parametrize("symmetric", [True, False])
def test_bfs_predecessors(mg, oracle, device, symmetric):
    g = mg.rmat_graph(14, 16, seed=11, symmetric=symmetric)
    src = int(np.argmax(g.degrees()))
    with mg.Graph(g, symmetric=symmetric, build_csc=not symmetric) as dg:
        d, p = dg.bfs_pred(src)
    assert np.array_equal(d, oracle.bfs(g.row_offsets, g.col_idx, src))
    assert np.array_equal(p, _pred_reference(g, d))
    # every reached vertex other than the source has a parent one level up
    reached = (d > 0)
    assert np.all(p[reached] >= 0) and np.all(d[p[reached]] == d[reached] - 1)
    assert np.all(p[d < 0] == -1)


def test_bfs_predecessors_fixtures(mg, oracle, device):
    for name in ["small.mtx", "chesapeake.mtx"]:
        g = load_mtx(mg, name)
        with mg.Graph(g, build_csc=True) as dg:
            for src in range(min(g.n, 4)):
                d, p = dg.bfs_pred(src)
                assert np.array_equal(p, _pred_reference(g, d)), (name, src)


# ---------------------------------------------------------------- uncaptured steps
_UNCAPTURED_CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
sys.path.insert(0, sys.argv[1] + "/tests")
import mirrorofmapgraph_amd as mg
import oracle
mg.check(mg.lib().mg_device_init(0))
rm = mg.rmat_graph(14, 16, seed=5, weight_max=64)
grid = mg.grid_graph(128, seed=6, weight_max=16)
for g in (rm, grid):
    with mg.Graph(g, symmetric=True) as dg:
        for s in (0, int(mg.pick_sources(g, 1, seed=3)[0])):
            assert np.array_equal(dg.bfs(s), oracle.bfs(g.row_offsets, g.col_idx, s)), ("bfs", s)
            assert np.array_equal(dg.sssp(s), oracle.sssp(g.row_offsets, g.col_idx, g.values, s)), ("sssp", s)
        assert np.array_equal(dg.cc(), oracle.cc(g.row_offsets, g.col_idx))
        srcs = mg.pick_sources(g, 40, seed=9)
        rows = dg.bfs_batch(srcs)
        srows = dg.sssp_batch(srcs)
        for j in (0, 17, 39):
            assert np.array_equal(rows[j], oracle.bfs(g.row_offsets, g.col_idx, int(srcs[j]))), ("bfs_batch", j)
            assert np.array_equal(srows[j], oracle.sssp(g.row_offsets, g.col_idx, g.values, int(srcs[j]))), j
gd = mg.rmat_graph(12, 16, seed=5, symmetric=False)
with mg.Graph(gd, build_csc=True) as dg:
    r = dg.pagerank(mg.options(pr_mode=mg.MG_PR_ALL, iter_num=20))
exp, _, _ = oracle.pr_jacobi_f32(gd.row_offsets, gd.col_idx, iters=20)
assert float(np.max(np.abs(r - exp))) / gd.n <= 1e-6
print("uncaptured ok", flush=True)
"""


def test_uncaptured_steps_match_oracle(mg, device):
    """MG_NO_GRAPH=1 launches every step directly instead of replaying captured HIP
    graphs (the path the server also takes when capture fails); the results must be
    the same.  The switch is read once per process, so it runs in a child."""
    import os
    import subprocess
    import sys
    from pathlib import Path

    root = str(Path(__file__).resolve().parents[1])
    env = dict(os.environ, MG_NO_GRAPH="1")
    r = subprocess.run([sys.executable, "-c", _UNCAPTURED_CHILD, root], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and "uncaptured ok" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]


# ---------------------------------------------------------------- overflow and allocation failures
def test_queue_overflow_regrows_and_replays(mg, oracle, device):
    """A tiny max_queue_sizing makes the edge-chunk queue (BFS / SSSP levels) and the
    wave-task queue (batched levels) overflow; the engine regrows the queue 4x and replays
    the search until it fits, counting the regrows in mg_stats.overflow_regrows, and the
    grown size stays with the graph.  The directed graph has no in-edge view, so every
    level pushes through the queues."""
    g = mg.rmat_graph(16, 16, seed=42, symmetric=False, weight_max=64)
    tiny = mg.options(max_queue_sizing=1e-4)
    exp_b = oracle.bfs(g.row_offsets, g.col_idx, 0)
    exp_s = oracle.sssp(g.row_offsets, g.col_idx, g.values, 0)
    with mg.Graph(g) as dg:
        d, st = dg.bfs(0, tiny, stats=True)
        assert np.array_equal(d, exp_b) and st.overflow_regrows > 0
        d, st = dg.bfs(0, tiny, stats=True)
        assert np.array_equal(d, exp_b) and st.overflow_regrows == 0
        d, st = dg.bfs(0, stats=True)
        assert np.array_equal(d, exp_b) and st.overflow_regrows == 0
    with mg.Graph(g) as dg:
        s, st = dg.sssp(0, tiny, stats=True)
        assert np.array_equal(s, exp_s) and st.overflow_regrows > 0
    srcs = mg.pick_sources(g, 64, seed=1)
    with mg.Graph(g) as dg:
        rows, st = dg.bfs_batch(srcs, tiny, stats=True)
        assert st.overflow_regrows > 0
        srows, st2 = dg.sssp_batch(srcs, tiny, stats=True)
        for j in (0, 31, 63):
            assert np.array_equal(rows[j], oracle.bfs(g.row_offsets, g.col_idx, int(srcs[j]))), j
            assert np.array_equal(srows[j], oracle.sssp(g.row_offsets, g.col_idx, g.values, int(srcs[j]))), j


_ALLOC_FAIL_CHILD = r"""
import os, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import mirrorofmapgraph_amd as mg
import oracle
mg.check(mg.lib().mg_device_init(0))
g = mg.rmat_graph(11, 8, seed=3, weight_max=16)
srcs = mg.pick_sources(g, 8, seed=2)
exp = {"bfs": oracle.bfs(g.row_offsets, g.col_idx, 0), "sssp": oracle.sssp(g.row_offsets, g.col_idx, g.values, 0),
       "cc": oracle.cc(g.row_offsets, g.col_idx)}
pr_exp, _, _ = oracle.pr_jacobi_f32(g.row_offsets, g.col_idx, iters=10)
refused = 0
for k in range(1, 90):
    os.environ["MG_FAIL_ALLOC_AT"] = str(k)  # the k-th device allocation from now on is refused
    def attempt(fn):
        global refused
        try:
            return fn()
        except mg.MgError as e:
            assert e.code == mg.MG_ERR_OOM, (k, e)
            refused += 1
            return fn()  # the hook fires once: the same call now succeeds
    dg = attempt(lambda: mg.Graph(g, symmetric=True))
    assert np.array_equal(attempt(lambda: dg.bfs(0)), exp["bfs"]), k
    assert np.array_equal(attempt(lambda: dg.sssp(0)), exp["sssp"]), k
    assert np.array_equal(attempt(lambda: dg.cc()), exp["cc"]), k
    r = attempt(lambda: dg.pagerank(mg.options(pr_mode=mg.MG_PR_ALL, iter_num=10)))
    assert float(np.max(np.abs(r - pr_exp))) / g.n <= 1e-6, k
    rows = attempt(lambda: dg.bfs_batch(srcs))
    assert np.array_equal(rows[0], oracle.bfs(g.row_offsets, g.col_idx, int(srcs[0]))), k
    srows = attempt(lambda: dg.sssp_batch(srcs))
    assert np.array_equal(srows[1], oracle.sssp(g.row_offsets, g.col_idx, g.values, int(srcs[1]))), k
    attempt(lambda: dg.ppr_batch(srcs[:4]))
    dg.close()
del os.environ["MG_FAIL_ALLOC_AT"]
assert refused > 30, refused
print("alloc failures ok", refused, flush=True)
"""


def test_allocation_failures_are_status_codes(mg, device):
    """With the MG_FAIL_ALLOC_AT hook the k-th device allocation is refused, for every k
    across graph creation and every search kind: each call fails with MG_ERR_OOM (never
    a crash or a wrong result), leaves the graph usable, and the same call then succeeds.
    Runs in a child process so the hook cannot touch other tests."""
    import subprocess
    import sys
    from pathlib import Path

    root = str(Path(__file__).resolve().parents[1])
    r = subprocess.run([sys.executable, "-c", _ALLOC_FAIL_CHILD, root], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "alloc failures ok" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]


@pytest.mark.parametrize("delta,switch", [(0, 0), (1, 0), (2, -1), (4, 1 << 30), (3, 1000), (-1, 0)])
def test_sssp_bucketed_whole_gpu_levels(mg, oracle, device, delta, switch):
    """Whole-GPU SSSP levels relax only pending vertices below the bucket bound and carry
    the others to the next frontier; the bound is lifted once the pending out-edges fall
    below sssp_bucket_edges (-1: never; huge: at the first empty bucket).  Every width and
    switch point gives the oracle's distances, on a symmetric R-MAT (range push on wide
    levels), a directed R-MAT with its in-edge view (directed = 0 expands both views)
    and a grid."""
    o = mg.options(sssp_delta=delta, sssp_bucket_edges=switch)
    rm = mg.rmat_graph(16, 16, seed=13, weight_max=64)
    with mg.Graph(rm, symmetric=True) as dg:
        for src in (0, int(mg.pick_sources(rm, 1, seed=4)[0])):
            got, st = dg.sssp(src, o, stats=True)
            exp = oracle.sssp(rm.row_offsets, rm.col_idx, rm.values, src)
            assert np.array_equal(got, exp), (delta, switch, src, np.nonzero(got != exp)[0][:8])
            assert st.edges_traversed == oracle.teps_edges(rm.row_offsets, exp, INT_MAX)
    gd = mg.rmat_graph(14, 16, seed=14, symmetric=False, weight_max=64)
    with mg.Graph(gd, build_csc=True) as dg:
        for directed in (1, 0):
            got = dg.sssp(0, mg.options(sssp_delta=delta, sssp_bucket_edges=switch, directed=directed))
            exp = oracle.sssp(gd.row_offsets, gd.col_idx, gd.values, 0, directed=directed)
            assert np.array_equal(got, exp), (delta, switch, directed)
    grid = mg.grid_graph(96, seed=15, weight_max=16)
    with mg.Graph(grid, symmetric=True) as dg:
        assert np.array_equal(dg.sssp(0, o), oracle.sssp(grid.row_offsets, grid.col_idx, grid.values, 0))


# ---------------------------------------------------------------- degree-ordered twin
@pytest.mark.parametrize("which", ["rmat14", "grid", "random_dir", "fixture"])
def test_relabelled_twin_matches_oracle(mg, oracle, device, which):
    """A graph created with MG_GRAPH_RELABEL runs BFS / SSSP on its degree-ordered copy
    (layout 0) and maps sources and results back: bit-exact with the oracle and with the
    caller-numbered searches (layout 1), TEPS counts equal."""
    sym = which in ("rmat14", "grid")
    if which == "rmat14":
        g = mg.rmat_graph(14, 16, seed=42, weight_max=64)
    elif which == "grid":
        g = mg.grid_graph(96, seed=3, weight_max=16)
    elif which == "random_dir":
        g = random_graph(mg, 4000, 30000, 11, symmetric=False, wmax=20)
    else:
        g = load_mtx(mg, "chesapeake.mtx")
    srcs = [0, int(np.argmax(g.degrees())), g.n - 1, g.n // 2]
    with mg.Graph(g, symmetric=sym, build_csc=not sym, relabel=True) as dg:
        for src in srcs:
            exp_b = oracle.bfs(g.row_offsets, g.col_idx, src)
            exp_s = oracle.sssp(g.row_offsets, g.col_idx, g.values, src)
            for lay in (0, 1):
                o = mg.options(layout=lay)
                b, sb = dg.bfs(src, o, stats=True)
                s, ss = dg.sssp(src, o, stats=True)
                assert np.array_equal(b, exp_b), (which, src, lay)
                assert np.array_equal(s, exp_s), (which, src, lay)
                assert sb.edges_traversed == oracle.teps_edges(g.row_offsets, exp_b)
                assert ss.edges_traversed == oracle.teps_edges(g.row_offsets, exp_s, INT_MAX)
        # the calls that keep the caller's numbering are unaffected by the twin
        assert np.array_equal(dg.cc(), oracle.cc(g.row_offsets, g.col_idx))
        if sym:
            d, p = dg.bfs_pred(srcs[1])
            assert np.array_equal(d, oracle.bfs(g.row_offsets, g.col_idx, srcs[1]))


def test_relabelled_twin_device_output_and_bytes(mg, oracle, device):
    """Results of a twin search written to a device buffer (out_on_device) are in the
    caller's numbering too; the twin's topology shows in the graph's device bytes."""
    import ctypes as C

    g = mg.rmat_graph(12, 16, seed=9, weight_max=64)
    with mg.Graph(g, symmetric=True) as plain, mg.Graph(g, symmetric=True, relabel=True) as dg:
        assert dg.device_bytes() >= plain.device_bytes() + 4 * (g.n + 1 + 2 * g.m)
        hip = C.CDLL("libamdhip64.so")
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), C.c_size_t(4 * g.n)) == 0
        try:
            out = np.empty(g.n, np.int32)
            mg.check(mg.lib().mg_sssp(dg.handle, 5, p, C.byref(mg.options(out_on_device=1)), None))
            assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(4 * g.n), 2) == 0
        finally:
            hip.hipFree(p)
        assert np.array_equal(out, oracle.sssp(g.row_offsets, g.col_idx, g.values, 5))


@pytest.mark.parametrize("mode", ["all", "gas"])
def test_pagerank_on_relabelled_twin(mg, oracle, device, mode):
    """PageRank on the degree-ordered copy: ranks mapped back to the caller's numbering,
    within the FP32 tolerance of the oracle (in-edges are summed in the copy's order)."""
    g = mg.rmat_graph(14, 16, seed=42, symmetric=False)
    with mg.Graph(g, build_csc=True, relabel=True) as dg:
        if mode == "all":
            got, st = dg.pagerank(mg.options(pr_mode=mg.MG_PR_ALL, iter_num=50), stats=True)
            assert st.iterations == 50
            exp, _, _ = oracle.pr_jacobi_f32(g.row_offsets, g.col_idx, iters=50)
        else:
            got, st = dg.pagerank(stats=True)
            exp, it = oracle.pr_gas(g.row_offsets, g.col_idx)
            assert abs(st.iterations - it) <= 1
        same = dg.pagerank(mg.options(pr_mode=mg.MG_PR_ALL, iter_num=50, layout=1)) if mode == "all" else None
    assert _pr_err(got, exp, g.n) <= PR_TOL
    if same is not None:
        assert _pr_err(same, exp, g.n) <= PR_TOL


# ---------------------------------------------------------------- host threads
def test_concurrent_host_threads(mg, oracle, device):
    """Searches from several host threads at once: each thread on its own graph, and all
    of them on one shared graph (calls on a handle are serialised by its mutex; each
    graph has its own stream).  Every result equals the oracle's."""
    import threading

    a = mg.rmat_graph(13, 16, seed=31, weight_max=64)
    b = mg.grid_graph(64, seed=32, weight_max=16)
    shared = mg.rmat_graph(12, 16, seed=33, weight_max=64)
    exp = {}
    for name, g in (("a", a), ("b", b), ("shared", shared)):
        for s in (0, 7, 100):
            exp[name, "bfs", s] = oracle.bfs(g.row_offsets, g.col_idx, s)
            exp[name, "sssp", s] = oracle.sssp(g.row_offsets, g.col_idx, g.values, s)
    dgs = {"a": mg.Graph(a, symmetric=True), "b": mg.Graph(b, symmetric=True, relabel=True),
           "shared": mg.Graph(shared, symmetric=True)}
    errors = []

    def worker(own):
        try:
            for rep in range(6):
                for name in (own, "shared"):
                    for s in (0, 7, 100):
                        dg = dgs[name]
                        if not np.array_equal(dg.bfs(s), exp[name, "bfs", s]):
                            errors.append((own, name, "bfs", s, rep))
                        if not np.array_equal(dg.sssp(s), exp[name, "sssp", s]):
                            errors.append((own, name, "sssp", s, rep))
        except Exception as e:  # reported below, in the main thread
            errors.append((own, repr(e)))

    try:
        th = [threading.Thread(target=worker, args=(o,)) for o in ("a", "b", "a", "b")]
        for t in th:
            t.start()
        for t in th:
            t.join()
    finally:
        for d in dgs.values():
            d.close()
    assert not errors, errors[:5]
