"""Batched multi-source BFS and SSSP on the GPU: row s of mg_bfs_batch /
mg_sssp_batch must equal a single-source search from srcs[s], bit for bit, and the
stats must be the sums of the per-source counts.  Small cases are checked against the CPU oracle, the
R-MAT scale-20 case (256 sources, the batched workload's size) against mg_bfs on
a sample of rows plus size-independent properties of every row."""
import numpy as np
import pytest

from conftest import load_mtx, random_graph

pytestmark = pytest.mark.gpu


def _oracle_rows(oracle, g, srcs, directed=1):
    return np.stack([oracle.bfs(g.row_offsets, g.col_idx, int(s), directed=directed) for s in srcs])


def _check_stats(oracle, g, rows, st):
    assert st.edges_traversed == sum(oracle.teps_edges(g.row_offsets, r) for r in rows)
    assert st.vertices_reached == int((rows >= 0).sum())


@pytest.mark.parametrize("k", [1, 31, 64, 200, 256])
def test_batch_rmat_symmetric(mg, oracle, device, k):
    g = mg.rmat_graph(13, 16, seed=7)
    srcs = mg.pick_sources(g, k, seed=k)
    with mg.Graph(g, symmetric=True) as dg:
        got, st = dg.bfs_batch(srcs, stats=True)
    exp = _oracle_rows(oracle, g, srcs)
    bad = np.argwhere(got != exp)
    assert bad.size == 0, f"first mismatches (row, vertex): {bad[:5].tolist()}"
    _check_stats(oracle, g, exp, st)


@pytest.mark.parametrize("build_csc", [True, False])
def test_batch_directed(mg, oracle, device, build_csc):
    # directed R-MAT: pull levels read the CSC; without it every level pushes
    g = mg.rmat_graph(12, 8, seed=3, symmetric=False)
    srcs = mg.pick_sources(g, 96, seed=5)
    with mg.Graph(g, build_csc=build_csc) as dg:
        got, st = dg.bfs_batch(srcs, stats=True)
    exp = _oracle_rows(oracle, g, srcs)
    assert np.array_equal(got, exp)
    _check_stats(oracle, g, exp, st)


def test_batch_grid_high_diameter(mg, oracle, device):
    g = mg.grid_graph(96, seed=11)
    rng = np.random.default_rng(0)
    srcs = rng.integers(0, g.n, 77).astype(np.int32)
    srcs[5] = srcs[6]  # a repeated source gets identical rows
    with mg.Graph(g, symmetric=True) as dg:
        got, st = dg.bfs_batch(srcs, stats=True)
    exp = _oracle_rows(oracle, g, srcs)
    assert np.array_equal(got, exp)
    assert np.array_equal(got[5], got[6])
    _check_stats(oracle, g, exp, st)


def test_batch_fixtures_and_isolated_sources(mg, oracle, device):
    for name in ["small.mtx", "trivialGraph.mtx", "chesapeake.mtx"]:
        g = load_mtx(mg, name)
        srcs = np.arange(g.n, dtype=np.int32)[:256]
        with mg.Graph(g, build_csc=True) as dg:
            got = dg.bfs_batch(srcs)
        assert np.array_equal(got, _oracle_rows(oracle, g, srcs)), name
    # vertices without edges: the row holds only the source
    g = random_graph(mg, 3000, 2000, 4)
    iso = np.nonzero(g.degrees() == 0)[0][:40].astype(np.int32)
    srcs = np.concatenate([iso, mg.pick_sources(g, 40, seed=1)])
    with mg.Graph(g, symmetric=True) as dg:
        got = dg.bfs_batch(srcs)
    assert np.array_equal(got, _oracle_rows(oracle, g, srcs))


def test_batch_matches_single_source_and_iter_num(mg, device):
    g = mg.rmat_graph(14, 16, seed=9)
    srcs = mg.pick_sources(g, 130, seed=2)
    with mg.Graph(g, symmetric=True) as dg:
        for it in (0, 1, 3):
            opt = mg.options(iter_num=it)
            got = dg.bfs_batch(srcs, opt)
            for s in (0, 64, 129):
                assert np.array_equal(got[s], dg.bfs(int(srcs[s]), opt)), (it, s)
        # forced push-only and forced early pull give the same rows
        base = dg.bfs_batch(srcs)
        assert np.array_equal(base, dg.bfs_batch(srcs, mg.options(batch_alpha=1 << 30, batch_beta=1 << 30)))
        assert np.array_equal(base, dg.bfs_batch(srcs, mg.options(batch_alpha=1 << 20, batch_beta=1)))
        # a second, smaller call on the same graph reuses the state
        assert np.array_equal(base[:3], dg.bfs_batch(srcs[:3]))


def test_batch_rejects_bad_arguments(mg, device):
    g = mg.rmat_graph(10, 8, seed=1)
    with mg.Graph(g, symmetric=True) as dg:
        for srcs in ([], np.zeros(mg.MG_BATCH_MAX + 1, np.int32), [0, g.n], [-1]):
            with pytest.raises(mg.MgError):
                dg.bfs_batch(srcs)
    gd = mg.rmat_graph(10, 8, seed=1, symmetric=False)
    with mg.Graph(gd, build_csc=True) as dg:
        with pytest.raises(mg.MgError):
            dg.bfs_batch([0], mg.options(directed=0))


def test_batch_rmat20_256_sources(mg, device):
    g = mg.rmat_graph(20, 16, seed=42)
    srcs = mg.pick_sources(g, 256, seed=42)
    with mg.Graph(g, symmetric=True) as dg:
        got, st = dg.bfs_batch(srcs, stats=True)
        for s in (0, 1, 77, 128, 255):
            assert np.array_equal(got[s], dg.bfs(int(srcs[s]))), s
    # Sampled rows are exact BFS distances: the source is at 0, neighbours differ by
    # at most 1 (symmetric graph), components are reached whole, and every vertex at
    # depth d > 0 has a neighbour at d - 1.  Together these fix the distances.
    ro, ci = g.row_offsets, g.col_idx
    src_rows = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(ro))
    nz = np.diff(ro) > 0
    assert np.all(got[np.arange(256), srcs] == 0)
    for s in range(0, 256, 16):
        d = got[s].astype(np.int64)
        du, dv = d[src_rows], d[ci]
        both = (du >= 0) & (dv >= 0)
        assert np.all(np.abs(du[both] - dv[both]) <= 1)
        assert np.all((du >= 0) == (dv >= 0))
        mn = np.full(g.n, 1 << 40)
        mn[nz] = np.minimum.reduceat(np.where(dv >= 0, dv, 1 << 40), ro[:-1][nz])
        assert np.all(mn[d > 0] == d[d > 0] - 1)
    assert st.vertices_reached == int((got >= 0).sum())
    deg = np.diff(ro).astype(np.int64)
    assert st.edges_traversed == sum(int(deg[row >= 0].sum()) for row in got)


# ---------------------------------------------------------------- batched SSSP
INT_MAX = 2**31 - 1


def _oracle_sssp_rows(oracle, g, srcs):
    return np.stack([oracle.sssp(g.row_offsets, g.col_idx, g.values, int(s)) for s in srcs])


def _check_sssp_stats(g, rows, st):
    deg = np.diff(g.row_offsets).astype(np.int64)
    assert st.vertices_reached == int((rows != INT_MAX).sum())
    assert st.edges_traversed == sum(int(deg[r != INT_MAX].sum()) for r in rows)


@pytest.mark.parametrize("k", [1, 17, 64, 256])
def test_sssp_batch_rmat(mg, oracle, device, k):
    g = mg.rmat_graph(13, 16, seed=7, weight_max=64)
    srcs = mg.pick_sources(g, k, seed=k + 100)
    with mg.Graph(g, symmetric=True) as dg:
        got, st = dg.sssp_batch(srcs, stats=True)
    exp = _oracle_sssp_rows(oracle, g, srcs)
    bad = np.argwhere(got != exp)
    assert bad.size == 0, f"first mismatches (row, vertex): {bad[:5].tolist()}"
    _check_sssp_stats(g, exp, st)


@pytest.mark.parametrize("delta", [0, -1, 1, 5, 300])
def test_sssp_batch_directed_and_grid(mg, oracle, device, delta):
    # delta: bucket width (0 = automatic, -1 = unbucketed Bellman-Ford)
    opt = mg.options(sssp_delta=delta)
    g = mg.rmat_graph(12, 8, seed=3, symmetric=False, weight_max=1000)
    srcs = mg.pick_sources(g, 100, seed=9)
    with mg.Graph(g) as dg:
        got, st = dg.sssp_batch(srcs, opt, stats=True)
    exp = _oracle_sssp_rows(oracle, g, srcs)
    assert np.array_equal(got, exp)
    _check_sssp_stats(g, exp, st)
    g = mg.grid_graph(80, seed=5, weight_max=16)
    srcs = np.random.default_rng(1).integers(0, g.n, 45).astype(np.int32)
    srcs[1] = srcs[0]
    with mg.Graph(g, symmetric=True) as dg:
        got = dg.sssp_batch(srcs, opt)
    assert np.array_equal(got, _oracle_sssp_rows(oracle, g, srcs))


def test_sssp_batch_fixtures_zero_weights_and_errors(mg, oracle, device):
    for name in ["small.mtx", "chesapeake.mtx"]:
        g = load_mtx(mg, name)
        srcs = np.arange(g.n, dtype=np.int32)
        with mg.Graph(g) as dg:
            assert np.array_equal(dg.sssp_batch(srcs), _oracle_sssp_rows(oracle, g, srcs)), name
    g = random_graph(mg, 2000, 12000, 8, wmax=3)
    g.values = (g.values - 1).astype(np.int32)  # weights 0..2, zero-weight edges included
    srcs = mg.pick_sources(g, 70, seed=3)
    with mg.Graph(g, symmetric=True) as dg:
        assert np.array_equal(dg.sssp_batch(srcs), _oracle_sssp_rows(oracle, g, srcs))
        with pytest.raises(mg.MgError):
            dg.sssp_batch([g.n])
    gu = mg.rmat_graph(10, 8, seed=1)  # no weights
    with mg.Graph(gu, symmetric=True) as dg:
        with pytest.raises(mg.MgError):
            dg.sssp_batch([0])


def test_sssp_batch_rmat20_256_sources(mg, device):
    g = mg.rmat_graph(20, 16, seed=42, weight_max=64)
    srcs = mg.pick_sources(g, 256, seed=43)
    with mg.Graph(g, symmetric=True) as dg:
        got, st = dg.sssp_batch(srcs, stats=True)
        for s in (0, 100, 255):
            assert np.array_equal(got[s], dg.sssp(int(srcs[s]))), s
    # Sampled rows are exact shortest distances: d[src] = 0, no edge can shorten
    # any distance (d[u] <= d[v] + w), and every reached vertex other than the
    # source has a tight in-edge (d[u] = d[v] + w for some v) -- with positive
    # weights these conditions determine the distances.
    ro, ci, w = g.row_offsets, g.col_idx, g.values.astype(np.int64)
    src_rows = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(ro))
    assert np.all(got[np.arange(256), srcs] == 0)
    for s in range(0, 256, 32):
        d = got[s].astype(np.int64)
        dv, du = d[src_rows], d[ci]
        ok = dv != INT_MAX
        assert np.all(du[ok] <= dv[ok] + w[ok])
        tight = np.zeros(g.n, bool)
        tight[ci[ok & (du == dv + w)]] = True
        reached = d != INT_MAX
        reached[srcs[s]] = False
        assert np.all(tight[reached])
    _check_sssp_stats(g, got, st)


# ---------------------------------------------------------------- batched personalized PageRank
PPR_TOL = 1e-6  # max-abs against the float64 oracle (ranks sum to at most 1)


@pytest.mark.parametrize("symmetric", [True, False])
def test_ppr_batch_against_float64(mg, oracle, device, symmetric):
    g = mg.rmat_graph(12, 16, seed=5, symmetric=symmetric)
    srcs = mg.pick_sources(g, 64, seed=8)
    with mg.Graph(g, symmetric=symmetric, build_csc=not symmetric) as dg:
        got, st = dg.ppr_batch(srcs, stats=True)
    assert st.iterations == 20
    for j in range(0, 64, 7):
        exp = oracle.ppr(g.row_offsets, g.col_idx, int(srcs[j]), iters=20, teleport=0.15)
        assert np.max(np.abs(got[j].astype(np.float64) - exp)) <= PPR_TOL, j


@pytest.mark.parametrize("symmetric", [True, False])
def test_ppr_batch_on_relabelled_twin(mg, oracle, device, symmetric):
    # a graph with a degree-ordered twin runs the batch there (renamed sources, rows
    # mapped back): within the tolerance of the oracle, and the same bits alone
    g = mg.rmat_graph(12, 16, seed=15, symmetric=symmetric)
    srcs = mg.pick_sources(g, 24, seed=16)
    with mg.Graph(g, symmetric=symmetric, build_csc=not symmetric, relabel=True) as dg:
        got = dg.ppr_batch(srcs)
        alone = dg.ppr_batch(srcs[5:6])
    assert np.array_equal(got[5].view(np.int32), alone[0].view(np.int32))
    for j in range(0, 24, 5):
        exp = oracle.ppr(g.row_offsets, g.col_idx, int(srcs[j]), iters=20, teleport=0.15)
        assert np.max(np.abs(got[j].astype(np.float64) - exp)) <= PPR_TOL, j


def test_ppr_batch_rows_do_not_depend_on_companions(mg, device):
    # the same source gives the same bits alone, in a batch, and in another slot, at
    # every row width (8 / 16 / 32 / 64 columns for batches of 1-8 / 9-16 / 17-32 / 33-64)
    g = mg.rmat_graph(13, 16, seed=2)
    srcs = mg.pick_sources(g, 40, seed=4)
    with mg.Graph(g, symmetric=True) as dg:
        full = dg.ppr_batch(srcs)
        for j in (0, 17, 39):
            alone = dg.ppr_batch(srcs[j:j + 1])
            assert np.array_equal(full[j].view(np.int32), alone[0].view(np.int32)), j
        for k in (9, 20):
            part = dg.ppr_batch(srcs[:k])
            assert np.array_equal(full[:k].view(np.int32), part.view(np.int32)), k
        rev = dg.ppr_batch(srcs[::-1].copy())
        assert np.array_equal(full.view(np.int32), rev[::-1].view(np.int32))
        # other iteration counts and teleport values
        o = mg.options(ppr_iters=3, pr_teleport=0.3)
        x = dg.ppr_batch(srcs[:5], o)
    import oracle as orc
    for j in range(5):
        exp = orc.ppr(g.row_offsets, g.col_idx, int(srcs[j]), iters=3, teleport=0.3)
        assert np.max(np.abs(x[j] - exp)) <= PPR_TOL


# ---------------------------------------------------------------- several graphs in one batched search
def _min_reached(rows, unreached):
    """Elementwise minimum over rows, ignoring `unreached` entries (multi-source answer)."""
    big = np.where(rows == unreached, np.iinfo(np.int64).max, rows.astype(np.int64))
    m = big.min(0)
    return np.where(m == np.iinfo(np.int64).max, unreached, m).astype(np.int32)


def test_union_graph_runs_several_graphs_in_one_batched_search(mg, oracle, device):
    """Graphs of uneven size (R-MAT, a grid, a sparse random graph with isolated
    vertices) joined by mg_graph_union advance in one batched search: row j carries the
    j-th source of every part, and each part's segment of the row equals the
    single-graph batched search (and the oracle), BFS and SSSP."""
    hosts = [mg.rmat_graph(13, 16, seed=7, weight_max=64), mg.grid_graph(96, seed=11, weight_max=16),
             random_graph(mg, 3000, 2000, 4, wmax=9)]
    dgs = [mg.Graph(h, symmetric=True) for h in hosts]
    try:
        u = mg.Graph.union(dgs)
        assert u.n == sum(h.n for h in hosts) and u.m == sum(h.m for h in hosts)
        assert list(u.offsets) == [0, hosts[0].n, hosts[0].n + hosts[1].n, u.n]
        srcs = [mg.pick_sources(hosts[0], 40, seed=1), mg.pick_sources(hosts[1], 77, seed=2),
                np.array([0, 1, 2, 5, 7], np.int32)]
        with u:
            for kind in ("bfs", "sssp"):
                parts, st = u.parts_batch(kind, srcs, stats=True)
                te = 0
                for i, dg in enumerate(dgs):
                    sep, sst = (dg.bfs_batch if kind == "bfs" else dg.sssp_batch)(srcs[i], stats=True)
                    bad = np.argwhere(parts[i] != sep)
                    assert bad.size == 0, (kind, i, bad[:5].tolist())
                    te += sst.edges_traversed
                    for j in (0, len(srcs[i]) - 1):
                        s = int(srcs[i][j])
                        exp = (oracle.bfs(hosts[i].row_offsets, hosts[i].col_idx, s) if kind == "bfs" else
                               oracle.sssp(hosts[i].row_offsets, hosts[i].col_idx, hosts[i].values, s))
                        assert np.array_equal(parts[i][j], exp), (kind, i, j)
                assert st.edges_traversed == te
            # a set of several sources in one part is a multi-source search
            g0 = hosts[0]
            two = [int(x) for x in mg.pick_sources(g0, 2, seed=9)]
            row = u.bfs_batch_sets([0, 0], two, 1)[0][: g0.n]
            single = np.stack([oracle.bfs(g0.row_offsets, g0.col_idx, s) for s in two])
            assert np.array_equal(row, _min_reached(single, -1))
            row = u.sssp_batch_sets([0, 0], two, 1)[0][: g0.n]
            single = np.stack([oracle.sssp(g0.row_offsets, g0.col_idx, g0.values, s) for s in two])
            assert np.array_equal(row, _min_reached(single, INT_MAX))
    finally:
        for d in dgs:
            d.close()


def test_union_graph_and_sets_argument_checks(mg, device):
    a = mg.Graph(mg.rmat_graph(8, 4, seed=1), symmetric=True)
    b = mg.Graph(mg.rmat_graph(8, 4, seed=2, symmetric=False))  # directed, unweighted
    try:
        with mg.Graph.union([a, b]) as u:
            assert not (u.flags & mg.MG_GRAPH_SYMMETRIC) and not (u.flags & mg.MG_GRAPH_WEIGHTED)
            with pytest.raises(mg.MgError):
                u.sssp_batch_sets([0], [0], 1)          # no weights
            for rows, srcs, k in (([0], [0], 2), ([2], [0], 2), ([0], [u.n], 1), ([], [], 1), ([0], [0], 300)):
                with pytest.raises(mg.MgError) as e:
                    u.bfs_batch_sets(rows, srcs, k)
                assert e.value.code == mg.MG_ERR_INVALID_ARG
            # a directed union pushes only; rows still match the separate searches
            got = u.parts_batch("bfs", [[0, 3], [1]])
            assert np.array_equal(got[0], a.bfs_batch([0, 3])) and np.array_equal(got[1], b.bfs_batch([1]))
        with pytest.raises(mg.MgError):
            mg.Graph.union([])
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("kind", ["bfs", "sssp"])
def test_batch_on_relabelled_twin(mg, oracle, device, kind):
    """A graph created with MG_GRAPH_RELABEL runs batched searches on its degree-ordered
    copy (sources renamed in, every row mapped back): rows equal the oracle's and the
    caller-numbered batch (layout 1), stats equal; the same through the per-row source
    sets entry point."""
    g = mg.rmat_graph(13, 16, seed=17, weight_max=64)
    srcs = mg.pick_sources(g, 96, seed=5)
    with mg.Graph(g, symmetric=True, relabel=True) as dg:
        fn = dg.bfs_batch if kind == "bfs" else dg.sssp_batch
        got, st = fn(srcs, stats=True)
        same, st1 = fn(srcs, mg.options(layout=1), stats=True)
        sets = dg.bfs_batch_sets if kind == "bfs" else dg.sssp_batch_sets
        rows_of = np.arange(96, dtype=np.int32) % 48
        two = sets(rows_of, srcs, 48)
    assert np.array_equal(got, same)
    assert st.edges_traversed == st1.edges_traversed and st.vertices_reached == st1.vertices_reached
    for j in (0, 5, 95):
        s = int(srcs[j])
        exp = oracle.bfs(g.row_offsets, g.col_idx, s) if kind == "bfs" else \
            oracle.sssp(g.row_offsets, g.col_idx, g.values, s)
        assert np.array_equal(got[j], exp), j
    # rows of the sets call: row r is the elementwise minimum over its two sources' rows
    for r in (0, 47):
        a, b = got[r], got[r + 48]
        if kind == "bfs":
            exp = np.where((a >= 0) & (b >= 0), np.minimum(a, b), np.maximum(a, b))
        else:
            exp = np.minimum(a, b)
        assert np.array_equal(two[r], exp), r
