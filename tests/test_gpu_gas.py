"""GPU parity of the generic GAS enactor (include/mapgraph/enactor.h) driving the
ported MapGraph programs (include/mapgraph/programs/*.h), selected through the C-ABI
with engine = MG_ENGINE_GAS.  Same bar as the specialised paths: BFS/SSSP/CC
bit-exact, PageRank (the reference program) within 1e-6 in normalised units."""
import numpy as np
import pytest

from conftest import load_mtx, random_graph

pytestmark = pytest.mark.gpu


def gas(mg, **kw):
    return mg.options(engine=mg.MG_ENGINE_GAS, **kw)


@pytest.mark.parametrize("name", ["small.mtx", "trivialGraph.mtx", "chesapeake.mtx"])
def test_gas_fixtures(mg, oracle, device, name):
    g = load_mtx(mg, name)
    with mg.Graph(g, build_csc=True) as dg:
        for directed in (1, 0):
            for src in range(min(g.n, 4)):
                d = dg.bfs(src, gas(mg, directed=directed))
                assert np.array_equal(d, oracle.bfs(g.row_offsets, g.col_idx, src, directed=directed))
                s = dg.sssp(src, gas(mg, directed=directed))
                assert np.array_equal(s, oracle.sssp(g.row_offsets, g.col_idx, g.values, src, directed=directed))
        assert np.array_equal(dg.cc(gas(mg)), oracle.cc(g.row_offsets, g.col_idx))


@pytest.mark.parametrize("symmetric", [True, False])
def test_gas_random(mg, oracle, device, symmetric):
    g = random_graph(mg, 3000, 20000, 11, symmetric=symmetric, wmax=64)
    src = int(np.argmax(g.degrees()))
    with mg.Graph(g, build_csc=not symmetric, symmetric=symmetric) as dg:
        assert np.array_equal(dg.bfs(src, gas(mg)), oracle.bfs(g.row_offsets, g.col_idx, src))
        assert np.array_equal(dg.sssp(src, gas(mg)), oracle.sssp(g.row_offsets, g.col_idx, g.values, src))
        assert np.array_equal(dg.cc(gas(mg)), oracle.cc(g.row_offsets, g.col_idx))


def test_gas_rmat16(mg, oracle, device):
    g = mg.rmat_graph(16, 16, seed=42, weight_max=64)
    with mg.Graph(g, symmetric=True) as dg:
        d, st = dg.bfs(0, gas(mg), stats=True)
        exp = oracle.bfs(g.row_offsets, g.col_idx, 0)
        assert np.array_equal(d, exp)
        assert st.iterations == exp.max() + 1  # one iteration per level, the last finds nothing
        assert np.array_equal(dg.sssp(0, gas(mg)), oracle.sssp(g.row_offsets, g.col_idx, g.values, 0))
        assert np.array_equal(dg.cc(gas(mg)), oracle.cc(g.row_offsets, g.col_idx))
        # the same answers as the specialised kernels
        assert np.array_equal(dg.bfs(0), d)


def test_gas_iter_num(mg, oracle, device):
    g = mg.rmat_graph(12, 16, seed=3)
    full = oracle.bfs(g.row_offsets, g.col_idx, 0)
    with mg.Graph(g, symmetric=True) as dg:
        for k in (1, 2):
            got = dg.bfs(0, gas(mg, iter_num=k))
            assert np.array_equal(got, np.where(full <= k, full, -1))


def test_gas_pagerank_reference_program(mg, oracle, device):
    g = mg.rmat_graph(13, 16, seed=9, symmetric=False)
    with mg.Graph(g, build_csc=True) as dg:
        got, st = dg.pagerank(gas(mg), stats=True)
        fast = dg.pagerank()  # the specialised pull path, same program semantics
    exp, it = oracle.pr_gas(g.row_offsets, g.col_idx)
    err = float(np.max(np.abs(got.astype(np.float64) - exp))) / g.n
    assert err <= 1e-6
    assert abs(st.iterations - it) <= 1
    assert float(np.max(np.abs(fast.astype(np.float64) - got))) / g.n <= 1e-6


def test_gas_errors(mg, device):
    g = mg.rmat_graph(10, 8, seed=1, symmetric=False)
    with mg.Graph(g) as dg:  # no in-edge view: PageRank gathers over in-edges
        with pytest.raises(mg.MgError) as e:
            dg.pagerank(gas(mg))
        assert e.value.code == mg.MG_ERR_UNSUPPORTED
        with pytest.raises(mg.MgError) as e:
            dg.bfs(0, mg.options(engine=7))
        assert e.value.code == mg.MG_ERR_INVALID_ARG
