"""Pins the CPU oracle (oracle/mg_oracle.c) before it is trusted as the checker.

* Known answers for the reference fixtures (smallRegressionGraphs), computed by hand
  from the files' edges (0-based, source 0).
* Independent library routines on random graphs: scipy.sparse.csgraph BFS order,
  Dijkstra and connected components.
* The two SSSP restatements against each other, and brute force on tiny graphs.
* PageRank: closed forms (2-cycle, star, chain), an exact float64 power iteration
  with numpy, the Gauss-Seidel fixed point, and hand-derived known answers that fix
  the program's activation rule (next frontier = out-neighbours of vertices that
  moved by >= eps, Jacobi within an iteration) and the in-place sweep order of CPUPR.
  Activating the changed vertices themselves, never deactivating, updating the
  program's ranks in place, or sweeping CPUPR in Jacobi order each fail one of them.
* The union-find CC checker against the CPUCC restatement and scipy.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

from conftest import load_mtx, random_graph

INT_MAX = 2**31 - 1


def to_scipy(g, weights=None):
    n = g.n
    data = np.ones(g.m) if weights is None else weights.astype(np.float64)
    rows = np.repeat(np.arange(n), np.diff(g.row_offsets))
    return sp.csr_matrix((data, (rows, g.col_idx)), shape=(n, n))


# ---------------------------------------------------------------- fixtures (KATs)
def test_small_known_answers(mg, oracle):
    g = load_mtx(mg, "small.mtx")
    assert g.n == 7 and g.m == 11
    assert oracle.bfs(g.row_offsets, g.col_idx, 0).tolist() == [0, 1, 1, 1, 2, 2, 2]
    assert oracle.sssp(g.row_offsets, g.col_idx, g.values, 0).tolist() == [0, 1, 4, 1, 2, 2, 2]
    assert oracle.sssp(g.row_offsets, g.col_idx, g.values, 0, quadratic=True).tolist() == [0, 1, 4, 1, 2, 2, 2]
    assert oracle.cc(g.row_offsets, g.col_idx).tolist() == [0] * 7


def test_trivial_known_answers(mg, oracle):
    g = load_mtx(mg, "trivialGraph.mtx")
    assert oracle.bfs(g.row_offsets, g.col_idx, 0).tolist() == [0, 1, 2, 3, 4]
    assert oracle.sssp(g.row_offsets, g.col_idx, g.values, 0).tolist() == [0, 1, 2, 3, 4]
    assert oracle.cc(g.row_offsets, g.col_idx).tolist() == [0] * 5
    # from the last vertex only the 4 <-> 3 (0-based) cycle is reachable
    assert oracle.bfs(g.row_offsets, g.col_idx, 4).tolist() == [-1, -1, -1, 1, 0]


def test_chesapeake_known_answers(mg, oracle):
    g = load_mtx(mg, "chesapeake.mtx")
    assert g.n == 39 and g.m == 170
    # lower triangle only: as loaded (directed), vertex 0 has no out-edges
    d = oracle.bfs(g.row_offsets, g.col_idx, 0, directed=1)
    assert d[0] == 0 and (d[1:] == -1).all()
    und = oracle.bfs(g.row_offsets, g.col_idx, 0, directed=0)
    expect = [0, 2, 2, 2, 2, 2, 1, 1, 2, 2, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2,
              2, 1, 1, 2, 1, 2, 1]
    assert und.tolist() == expect
    a = to_scipy(g)
    ref = cg.shortest_path(a + a.T, unweighted=True, indices=0)
    assert np.array_equal(und, ref.astype(np.int64))
    assert oracle.cc(g.row_offsets, g.col_idx).tolist() == [0] * 39


def test_sssp_unreached_convention(mg, oracle):
    g = load_mtx(mg, "trivialGraph.mtx")
    ref_cpu = oracle.sssp(g.row_offsets, g.col_idx, g.values, 4, unreached=100000000, quadratic=True)
    assert ref_cpu.tolist() == [100000000, 100000000, 100000000, 1, 0]  # sssp.cu:138-139
    gpu_conv = oracle.sssp(g.row_offsets, g.col_idx, g.values, 4)
    assert gpu_conv.tolist() == [INT_MAX, INT_MAX, INT_MAX, 1, 0]  # sssp.h:43


# ---------------------------------------------------------------- library routines
@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("symmetric", [True, False])
def test_bfs_matches_scipy(mg, oracle, seed, symmetric):
    g = random_graph(mg, 300, 700, seed, symmetric=symmetric)
    src = int(np.argmax(g.degrees()))
    d = oracle.bfs(g.row_offsets, g.col_idx, src)
    order, pred = cg.breadth_first_order(to_scipy(g), src, directed=True, return_predecessors=True)
    expect = np.full(g.n, -1)
    expect[src] = 0
    for v in order[1:]:
        expect[v] = expect[pred[v]] + 1
    assert np.array_equal(d, expect)


@pytest.mark.parametrize("seed", [4, 5])
@pytest.mark.parametrize("symmetric", [True, False])
def test_sssp_matches_scipy_dijkstra(mg, oracle, seed, symmetric):
    g = random_graph(mg, 400, 1500, seed, symmetric=symmetric, wmax=64)
    src = int(np.argmax(g.degrees()))
    d = oracle.sssp(g.row_offsets, g.col_idx, g.values, src)
    dq = oracle.sssp(g.row_offsets, g.col_idx, g.values, src, quadratic=True)
    ref = cg.dijkstra(to_scipy(g, g.values), directed=True, indices=src)
    expect = np.where(np.isinf(ref), INT_MAX, ref).astype(np.int64)
    assert np.array_equal(d, expect)
    assert np.array_equal(dq, expect)


def test_sssp_undirected_uses_edge_ids(mg, oracle):
    # directed=0 relaxes in-edges with the weight of the CSR edge they belong to
    g = random_graph(mg, 200, 500, 9, symmetric=False, wmax=16)
    d = oracle.sssp(g.row_offsets, g.col_idx, g.values, 0, directed=0)
    dq = oracle.sssp(g.row_offsets, g.col_idx, g.values, 0, directed=0, quadratic=True)
    a = to_scipy(g, g.values)
    ref = cg.dijkstra(a, directed=False, indices=0)
    assert np.array_equal(d, np.where(np.isinf(ref), INT_MAX, ref).astype(np.int64))
    assert np.array_equal(d, dq)


@pytest.mark.parametrize("seed", [6, 7])
@pytest.mark.parametrize("symmetric", [True, False])
def test_cc_matches_scipy(mg, oracle, seed, symmetric):
    g = random_graph(mg, 500, 400, seed, symmetric=symmetric)
    lab = oracle.cc(g.row_offsets, g.col_idx)
    ncomp, comp = cg.connected_components(to_scipy(g), directed=True, connection="weak")
    expect = np.empty(g.n, np.int64)
    for c in range(ncomp):
        members = np.nonzero(comp == c)[0]
        expect[members] = members.min()
    assert np.array_equal(lab, expect)


@pytest.mark.parametrize("seed", [6, 7, 8])
@pytest.mark.parametrize("symmetric", [True, False])
def test_cc_union_find_matches_cpucc(mg, oracle, seed, symmetric):
    # the union-find checker used at scale 22 against the CPUCC restatement and scipy
    g = random_graph(mg, 3000, 2500, seed, symmetric=symmetric)
    lab = oracle.cc_union_find(g.row_offsets, g.col_idx)
    assert np.array_equal(lab, oracle.cc(g.row_offsets, g.col_idx))
    ncomp, comp = cg.connected_components(to_scipy(g), directed=True, connection="weak")
    assert len(np.unique(lab)) == ncomp
    for name in ("small.mtx", "trivialGraph.mtx", "chesapeake.mtx"):
        f = load_mtx(mg, name)
        assert np.array_equal(oracle.cc_union_find(f.row_offsets, f.col_idx), oracle.cc(f.row_offsets, f.col_idx))


def test_cc_union_find_rmat(mg, oracle):
    g = mg.rmat_graph(14, 8, seed=4)
    assert np.array_equal(oracle.cc_union_find(g.row_offsets, g.col_idx), oracle.cc(g.row_offsets, g.col_idx))


def test_bfs_brute_force_tiny(mg, oracle):
    rng = np.random.default_rng(11)
    for trial in range(20):
        n = int(rng.integers(1, 9))
        g = random_graph(mg, n, int(rng.integers(0, 20)), 100 + trial, symmetric=bool(trial % 2))
        a = np.zeros((n, n), bool)
        for u in range(n):
            a[u, g.col_idx[g.row_offsets[u]:g.row_offsets[u + 1]]] = True
        src = int(rng.integers(0, n))
        expect = np.full(n, -1)
        reach = np.zeros(n, bool)
        reach[src] = True
        expect[src] = 0
        for k in range(1, n + 1):  # k-step reachability by boolean matrix powers
            nxt = (reach[:, None] & a).any(0) | reach
            expect[nxt & ~reach] = k
            reach = nxt
        assert np.array_equal(oracle.bfs(g.row_offsets, g.col_idx, src), expect)


def test_teps_count(mg, oracle):
    g = load_mtx(mg, "trivialGraph.mtx")
    d = oracle.bfs(g.row_offsets, g.col_idx, 4)
    assert oracle.teps_edges(g.row_offsets, d) == 2  # vertices 3 and 4, one out-edge each


# ---------------------------------------------------------------- PageRank
def _csr(edges, n):
    import mirrorofmapgraph_amd as mg
    s = np.array([e[0] for e in edges], np.int32)
    d = np.array([e[1] for e in edges], np.int32)
    return mg.csr_from_edges(n, s, d, mg.MG_CSR_SORT_COLS)


def test_pr_two_cycle_closed_form(oracle):
    g = _csr([(0, 1), (1, 0)], 2)
    # r = 0.15 + 0.85 r has the fixed point r = 1
    r64, _, _ = oracle.pr_jacobi_f64(g.row_offsets, g.col_idx, iters=400)
    assert np.allclose(r64, 1.0, atol=1e-12)
    r32, _, _ = oracle.pr_jacobi_f32(g.row_offsets, g.col_idx, iters=400)
    assert np.allclose(r32, 1.0, atol=1e-6)


def test_pr_star_closed_form(oracle):
    k = 5  # leaves 1..k point to hub 0; the hub has no out-edges (dangling)
    g = _csr([(i, 0) for i in range(1, k + 1)], k + 1)
    r, _, _ = oracle.pr_jacobi_f64(g.row_offsets, g.col_idx, iters=10)
    assert np.allclose(r[1:], 0.15)
    assert np.isclose(r[0], 0.15 + 0.85 * 0.15 * k)
    rg, it = oracle.pr_gas(g.row_offsets, g.col_idx)
    assert np.allclose(rg, r, atol=1e-6)


def test_pr_chain_exact_iterations(oracle):
    # 0 -> 1 -> 2: after t Jacobi iterations r2 = 0.15 + 0.85*r1_prev
    g = _csr([(0, 1), (1, 2)], 3)
    r, _, _ = oracle.pr_jacobi_f64(g.row_offsets, g.col_idx, iters=1)
    assert np.allclose(r, [0.15, 0.15 + 0.85 * 0.15, 0.15 + 0.85 * 0.15])
    r, _, _ = oracle.pr_jacobi_f64(g.row_offsets, g.col_idx, iters=2)
    assert np.allclose(r, [0.15, 0.2775, 0.15 + 0.85 * 0.2775])


def test_pr_f64_matches_numpy_power_iteration(mg, oracle):
    g = random_graph(mg, 300, 1500, 21, symmetric=False)
    a = to_scipy(g)  # a[u, v] = 1 for u -> v
    deg = np.asarray(a.sum(1)).ravel()
    inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    r = np.full(g.n, 0.15)
    for _ in range(30):
        r = 0.15 + 0.85 * (a.T @ (r * inv))
    got, it, _ = oracle.pr_jacobi_f64(g.row_offsets, g.col_idx, iters=30)
    assert it == 30
    assert np.allclose(got, r, rtol=0, atol=1e-12)
    got32, _, _ = oracle.pr_jacobi_f32(g.row_offsets, g.col_idx, iters=30)
    assert np.max(np.abs(got32 - r)) < 1e-4


def test_pr_gauss_seidel_fixed_point(mg, oracle):
    g = random_graph(mg, 200, 1000, 22, symmetric=False)
    r, sweeps = oracle.pr_gauss_seidel(g.row_offsets, g.col_idx)
    assert sweeps > 1
    # at the stop every vertex moved by < 0.01 in the last sweep, so r is close to
    # the float64 fixed point
    ref, _, _ = oracle.pr_jacobi_f64(g.row_offsets, g.col_idx, iters=500)
    assert np.max(np.abs(r - ref)) < 0.1


def test_pr_gauss_seidel_sweep_order(oracle):
    """CPUPR (PageRank.cu:136-156) updates ranks in place in vertex order, so a sweep
    already sees the new ranks of lower ids.  Hand-derived sweeps:
    chain 0 -> 1 -> 2: sweep 1 gives r1 = 0.15 + 0.85*0.15 = 0.2775 and then
    r2 = 0.15 + 0.85*0.2775 = 0.385875 (a Jacobi sweep would give r2 = 0.2775); sweep 2
    moves nothing, so 2 sweeps.
    chain 2 -> 1 -> 0: sweep 1 reads the old ranks of higher ids: r0 = r1 = 0.2775,
    r2 = 0.15; sweep 2: r0 = 0.385875 (moved by 0.108), sweep 3 moves nothing."""
    g = _csr([(0, 1), (1, 2)], 3)
    r1, s1 = oracle.pr_gauss_seidel(g.row_offsets, g.col_idx, max_sweeps=1)
    assert s1 == 1 and np.allclose(r1, [0.15, 0.2775, 0.385875], rtol=0, atol=1e-6)
    r, s = oracle.pr_gauss_seidel(g.row_offsets, g.col_idx)
    assert s == 2 and np.allclose(r, [0.15, 0.2775, 0.385875], rtol=0, atol=1e-6)
    g = _csr([(2, 1), (1, 0)], 3)
    r1, _ = oracle.pr_gauss_seidel(g.row_offsets, g.col_idx, max_sweeps=1)
    assert np.allclose(r1, [0.2775, 0.2775, 0.15], rtol=0, atol=1e-6)
    r, s = oracle.pr_gauss_seidel(g.row_offsets, g.col_idx)
    assert s == 3 and np.allclose(r, [0.385875, 0.2775, 0.15], rtol=0, atol=1e-6)


def _activation_graph():
    """a=0 -> z=1; z -> y=2 and z -> sinks 4..15 (out-degree 13); y -> x=3."""
    edges = [(0, 1), (1, 2), (2, 3)] + [(1, s) for s in range(4, 16)]
    return _csr(edges, 16)


def test_pr_gas_activation_known_answers(oracle):
    """The GPU PageRank program (PageRank.h:233-283; next frontier = out-neighbours of
    the vertices whose rank moved by >= eps; Jacobi within an iteration), derived by hand
    on _activation_graph with t = 0.15:
    iteration 1 (all active): z = 0.15 + 0.85*0.15 = 0.2775 (moved, changed);
      y = sinks = 0.15 + 0.85*0.15/13 = 0.1598077 (moved 0.0098 < 0.01: unchanged);
      x = 0.15 + 0.85*0.15 = 0.2775 (changed; no out-edges); a = 0.15 (unchanged).
      Next frontier = out(z) + out(x) = {y, sinks}.
    iteration 2: y = sinks = 0.15 + 0.85*0.2775/13 = 0.1681442 (moved 0.0083: unchanged);
      frontier empty: 2 iterations.  x keeps the stale 0.2775: y never moved by 0.01 at
      once, so x is never recomputed from y's new rank (a converged Jacobi run would give
      0.15 + 0.85*0.1681442).
    Activating the changed vertices themselves, never deactivating, or updating in place
    all give other ranks here."""
    g = _activation_graph()
    r, it = oracle.pr_gas(g.row_offsets, g.col_idx)
    y2 = 0.15 + 0.85 * 0.2775 / 13
    exp = np.array([0.15, 0.2775, y2, 0.2775] + [y2] * 12)
    assert it == 2
    assert np.allclose(r, exp, rtol=0, atol=1e-6), r
    r1, it1 = oracle.pr_gas(g.row_offsets, g.col_idx, iter_num=1)
    y1 = 0.15 + 0.85 * 0.15 / 13
    assert it1 == 1
    assert np.allclose(r1, [0.15, 0.2775, y1, 0.2775] + [y1] * 12, rtol=0, atol=1e-6)


def test_pr_gas_activation_threshold(oracle):
    """Same graph with eps = 0.005, by hand: iteration 1 as above but y and the sinks
    (moved 0.0098) are changed too, so x joins the frontier.  Iteration 2: y = sinks =
    0.1681442 (moved 0.0083, changed), x = 0.15 + 0.85*0.1598077 = 0.2858365 (moved
    0.0083, changed).  Iteration 3: frontier = out(y) = {x}: x = 0.15 + 0.85*0.1681442 =
    0.2929226 (moved 0.0071, changed; no out-edges).  Frontier empty: 3 iterations."""
    g = _activation_graph()
    r, it = oracle.pr_gas(g.row_offsets, g.col_idx, eps=0.005)
    y2 = 0.15 + 0.85 * 0.2775 / 13
    assert it == 3
    exp = np.array([0.15, 0.2775, y2, 0.15 + 0.85 * y2] + [y2] * 12)
    assert np.allclose(r, exp, rtol=0, atol=1e-6), r


def test_pr_gas_semantics(mg, oracle):
    g = random_graph(mg, 300, 1200, 23, symmetric=False)
    rg, it = oracle.pr_gas(g.row_offsets, g.col_idx)
    assert it > 1
    # every rank is within the activation threshold's reach of the fixed point
    ref, _, _ = oracle.pr_jacobi_f64(g.row_offsets, g.col_idx, iters=500)
    assert np.max(np.abs(rg - ref)) < 0.2
    # at the stop no vertex whose in-neighbours changed in the last iteration moves by
    # 0.01 or more: running once more from the final ranks leaves the frontier empty
    r_more, it_more = oracle.pr_gas(g.row_offsets, g.col_idx, iter_num=it + 5)
    assert it_more == it and np.array_equal(r_more, rg)
    # iter_num caps the iterations; one iteration equals one all-vertex Jacobi step
    r1, it1 = oracle.pr_gas(g.row_offsets, g.col_idx, iter_num=1)
    j1, _, _ = oracle.pr_jacobi_f32(g.row_offsets, g.col_idx, iters=1)
    assert it1 == 1 and np.array_equal(r1, j1)


# ---------------------------------------------------------------- personalized PageRank
def test_ppr_two_cycle_closed_form(oracle):
    # 0 <-> 1 from source 0: a' = t + (1-t) b, b' = (1-t) a
    ro, ci = np.array([0, 1, 2]), np.array([1, 0])
    t = 0.15
    a, b = 1.0, 0.0
    for k in range(1, 8):
        a, b = t + (1 - t) * b, (1 - t) * a
        x = oracle.ppr(ro, ci, 0, iters=k, teleport=t)
        assert abs(x[0] - a) < 1e-15 and abs(x[1] - b) < 1e-15


def test_ppr_star_closed_form(oracle):
    # centre 0, L leaves of degree 1, from the centre: c' = t + (1-t) L l, l' = (1-t) c / L
    L, t = 5, 0.2
    ro = np.concatenate([[0, L], L + np.arange(1, L + 1)])
    ci = np.concatenate([np.arange(1, L + 1), np.zeros(L, np.int64)])
    c, leaf = 1.0, 0.0
    for k in range(1, 6):
        c, leaf = t + (1 - t) * L * leaf, (1 - t) * c / L
        x = oracle.ppr(ro, ci, 0, iters=k, teleport=t)
        assert abs(x[0] - c) < 1e-14 and np.allclose(x[1:], leaf, atol=1e-14)


def test_ppr_mass_is_conserved_without_dangling_vertices(mg, oracle):
    g = mg.rmat_graph(9, 8, seed=3)
    src = int(np.argmax(g.degrees()))
    for k in (1, 7, 20):
        x = oracle.ppr(g.row_offsets, g.col_idx, src, iters=k)
        assert abs(x.sum() - 1.0) < 1e-12 and x.min() >= 0.0
