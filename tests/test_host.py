"""Host-side tests (no GPU): the C-ABI library loads and exports every declared
symbol, the generators match their numpy restatement bit for bit, the CSR/CSC
builders and the Matrix Market reader behave as the reference's loaders, and the
argument checks return status codes instead of crashing."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, load_mtx

HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
        for m in re.finditer(r"\b(mg_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_exports_every_declared_symbol(mg):
    L = mg.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    for n in mg.EXPORTED:
        assert n in names, f"{n} exported by the binding but not declared in include/"


def test_struct_layouts(mg):
    # mg_options / mg_stats sizes as laid out by the C compiler
    assert C.sizeof(mg.Options) == 4 * 10 + 8 + 4 * 2 + 4 * 8
    assert mg.Options.layout.offset == mg.Options.level_grid.offset + 4
    assert C.sizeof(mg.Stats) == 8 * 3 + 4 * 2 + 8 * 2 + 4 * 8
    o = mg.options()
    assert o.iter_num == 2**31 - 1 and o.threshold == 10000 and o.directed == 1
    assert o.pr_teleport == np.float32(0.15) and o.pr_active_eps == np.float32(0.01)
    assert C.sizeof(mg.ServerConfig) == 8 + 4 * 7 + 4 * 5
    assert mg.ServerConfig.thin_streams.offset == mg.ServerConfig.thin_lanes.offset + 4
    assert mg.ServerConfig.thin_slice_us.offset == mg.ServerConfig.thin_streams.offset + 4
    assert mg.ServerConfig.ppr_batch_min.offset == mg.ServerConfig.ppr_gate_us.offset + 4


def test_server_config_validation(mg):
    # an invalid configuration is refused before the server touches a device
    for bad in ({"max_slots": 0}, {"max_slots": 257}, {"steps_per_round": 65}, {"thin_streams": 9},
                {"thin_streams": -1}, {"thin_slice_us": -2}, {"ppr_gate_us": -2},
                {"ppr_batch_min": 65}, {"plain_kinds": 4}, {"ppr_iters": 0}):
        with pytest.raises(mg.MgError) as e:
            mg.Server(**bad)
        assert e.value.code == mg.MG_ERR_INVALID_ARG, bad
    with pytest.raises(KeyError):
        mg.Server(no_such_option=1)


def test_version_and_errors(mg):
    assert mg.lib().mg_version().decode().count(".") == 2
    with pytest.raises(mg.MgError) as e:
        mg.rmat_edges(0, 16)
    assert e.value.code == mg.MG_ERR_INVALID_ARG
    assert "mg_gen_rmat_edges" in mg.last_error()


def test_rmat_matches_numpy_restatement(mg):
    from oracle import gen_ref

    for scale, seed in [(6, 1), (11, 42)]:
        s, d = mg.rmat_edges(scale, 16, seed=seed)
        s2, d2 = gen_ref.rmat_edges(scale, 16, 0.57, 0.19, 0.19, seed)
        assert np.array_equal(s, s2) and np.array_equal(d, d2)
        assert s.min() >= 0 and s.max() < (1 << scale)


def test_rmat_quadrant_statistics(mg):
    # top-level quadrant frequencies follow (a, b, c, d)
    s, d = mg.rmat_edges(16, 16, seed=5)
    top_u = s >> 15
    top_v = d >> 15
    freq = np.array([np.mean((top_u == i) & (top_v == j)) for i, j in [(0, 0), (0, 1), (1, 0), (1, 1)]])
    assert np.allclose(freq, [0.57, 0.19, 0.19, 0.05], atol=0.003)


def test_csr_builder_matches_numpy(mg):
    from oracle import gen_ref

    s, d = mg.rmat_edges(12, 8, seed=3)
    for sym in (True, False):
        flags = mg.MG_CSR_DROP_LOOPS | mg.MG_CSR_DEDUPE | mg.MG_CSR_SORT_COLS | (mg.MG_CSR_SYMMETRIZE if sym else 0)
        g = mg.csr_from_edges(1 << 12, s, d, flags)
        ro, ci = gen_ref.csr(1 << 12, s, d, sym, True, True)
        assert np.array_equal(g.row_offsets, ro) and np.array_equal(g.col_idx, ci)
    gs = mg.rmat_graph(10, seed=9)
    a = set(zip(np.repeat(np.arange(gs.n), gs.degrees()).tolist(), gs.col_idx.tolist()))
    assert all((v, u) in a for (u, v) in a)  # symmetric
    assert all(u != v for (u, v) in a)  # no self-loops


def test_csr_from_coo_keeps_input_order(mg):
    # CsrGraph::FromCoo: stable sort by row only; duplicates and loops are kept
    s = np.array([2, 0, 2, 0, 1, 0], np.int32)
    d = np.array([1, 2, 0, 2, 1, 1], np.int32)
    v = np.array([10, 11, 12, 13, 14, 15], np.int32)
    g = mg.csr_from_edges(3, s, d, 0, vals=v)
    assert g.row_offsets.tolist() == [0, 3, 4, 6]
    assert g.col_idx.tolist() == [2, 2, 1, 1, 1, 0]
    assert g.values.tolist() == [11, 13, 15, 14, 10, 12]


def test_weights_symmetric_and_in_range(mg):
    from oracle import gen_ref

    g = mg.rmat_graph(10, seed=4, weight_max=64)
    assert g.values.min() >= 1 and g.values.max() <= 64
    assert np.array_equal(g.values, gen_ref.weights(g.row_offsets, g.col_idx, 5, 64))
    w = {}
    rows = np.repeat(np.arange(g.n), g.degrees())
    for u, v, x in zip(rows.tolist(), g.col_idx.tolist(), g.values.tolist()):
        w[(u, v)] = x
    assert all(w[(v, u)] == x for (u, v), x in w.items())


def test_grid_generator(mg):
    from oracle import gen_ref

    for W, H in [(5, 4), (64, 64)]:
        s, d = mg.grid_edges(W, H, 0.05, 7)
        s2, d2 = gen_ref.grid_edges(W, H, 0.05, 7)
        assert np.array_equal(s, s2) and np.array_equal(d, d2)
        E = H * (W - 1) + (H - 1) * W
        assert len(s) == E - int(np.floor(0.05 * E))
        assert ((d - s == 1) | (d - s == W)).all()
    s, d = mg.grid_edges(8, 8, 0.0, 1)
    assert len(s) == 2 * 8 * 7


def test_pick_sources(mg):
    from oracle import gen_ref

    g = mg.rmat_graph(12, seed=6)
    src = mg.pick_sources(g, 64, 99)
    assert np.array_equal(src, gen_ref.sources(g.row_offsets, 64, 99))
    assert (g.degrees()[src] > 0).all()


def test_csc_builder(mg):
    g = mg.rmat_graph(10, seed=8, symmetric=False)
    co, ri, ce = g.csc()
    assert co[0] == 0 and co[-1] == g.m
    rows = np.repeat(np.arange(g.n, dtype=np.int32), g.degrees())
    order = np.argsort(g.col_idx, kind="stable")
    assert np.array_equal(ce, order)
    assert np.array_equal(ri, rows[order])
    assert np.array_equal(g.col_idx[ce], np.repeat(np.arange(g.n), np.diff(co)))


@pytest.mark.parametrize("kind", ["rmat_sym", "rmat_dir", "multi"])
def test_degree_relabel(mg, kind):
    """mg_csr_relabel_degree: new ids rank vertices by out-degree (highest first, ties by
    id), every edge (u, v, w) appears as (new[u], new[v], w), rows are sorted; checked
    against a numpy restatement (stable argsort of -degree, lexicographic edge sort)."""
    if kind == "multi":  # duplicate edges and self-loops survive the renaming
        rng = np.random.default_rng(3)
        n = 300
        s, d = rng.integers(0, n, 4000).astype(np.int32), rng.integers(0, n, 4000).astype(np.int32)
        g = mg.csr_from_edges(n, s, d, 0, vals=rng.integers(1, 99, 4000).astype(np.int32))
    else:
        g = mg.rmat_graph(12, seed=5, symmetric=kind == "rmat_sym", weight_max=64)
    h, perm = mg.relabel_degree(g)
    deg = g.degrees().astype(np.int64)
    order = np.argsort(-deg, kind="stable")
    new = np.empty(g.n, np.int64)
    new[order] = np.arange(g.n)
    assert np.array_equal(perm, new)
    assert np.all(np.diff(h.degrees()) <= 0)  # degrees non-increasing in new ids
    assert np.array_equal(h.degrees(), deg[order])
    src = np.repeat(np.arange(g.n), deg)
    key = (new[src] << 32) | new[g.col_idx]
    o = np.argsort(key, kind="stable")
    assert np.array_equal(h.col_idx, (key[o] & 0xFFFFFFFF).astype(np.int32))
    assert np.array_equal(h.values, g.values[o])
    assert np.array_equal(np.repeat(np.arange(g.n), h.degrees()), (key[o] >> 32))


def test_mtx_reader(mg, tmp_path):
    n, s, d, v, sym = mg.read_mtx(GOLDEN / "small.mtx")
    assert n == 7 and len(s) == 11 and not sym
    assert (s[0], d[0], v[1]) == (0, 1, 4)
    n, s, d, v, sym = mg.read_mtx(GOLDEN / "chesapeake.mtx")
    assert n == 39 and len(s) == 170 and sym and (v == 1).all()  # pattern: weight 1
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate pattern general\n3 4 2\n1 2\n")
    with pytest.raises(mg.MgError) as e:
        mg.read_mtx(bad)
    assert e.value.code == mg.MG_ERR_FORMAT
    short = tmp_path / "short.mtx"
    short.write_text("3 3 2\n1 2\n")
    with pytest.raises(mg.MgError) as e:
        mg.read_mtx(short)
    assert e.value.code == mg.MG_ERR_FORMAT
    oob = tmp_path / "oob.mtx"
    oob.write_text("3 3 1\n1 4\n")
    with pytest.raises(mg.MgError):
        mg.read_mtx(oob)
    with pytest.raises(mg.MgError) as e:
        mg.read_mtx(tmp_path / "missing.mtx")
    assert e.value.code == mg.MG_ERR_IO


def test_graph_create_argument_checks(mg):
    L = mg.lib()
    h = C.c_void_p()
    ro = np.array([0, 1, 2], np.int32)
    col = np.array([1, 5], np.int32)  # 5 is outside [0, 2)
    rc = L.mg_graph_create(mg._p(ro), mg._p(col), None, 2, 2, 0, C.byref(h))
    assert rc == mg.MG_ERR_INVALID_ARG and not h.value
    ro_bad = np.array([0, 2, 1], np.int32)
    rc = L.mg_graph_create(mg._p(ro_bad), mg._p(np.array([0], np.int32)), None, 2, 1, 0, C.byref(h))
    assert rc == mg.MG_ERR_INVALID_ARG
    rc = L.mg_graph_create(mg._p(ro), mg._p(np.array([1, 0], np.int32)), None, 2, 2, 0x80, C.byref(h))
    assert rc == mg.MG_ERR_INVALID_ARG
    rc = L.mg_graph_create(mg._p(ro), mg._p(np.array([1, 0], np.int32)), None, 2, 2, mg.MG_GRAPH_WEIGHTED, C.byref(h))
    assert rc == mg.MG_ERR_INVALID_ARG
    # null graph handles are rejected, not dereferenced
    assert L.mg_bfs(None, 0, None, None, None) == mg.MG_ERR_INVALID_ARG
    assert L.mg_cc(None, None, None, None) == mg.MG_ERR_INVALID_ARG
    assert L.mg_graph_destroy(None) == mg.MG_OK


def test_fixture_loads_as_reference(mg):
    g = load_mtx(mg, "small.mtx")
    assert g.row_offsets.tolist() == [0, 3, 4, 7, 9, 10, 11, 11]
    assert g.values.tolist()[:3] == [1, 4, 1]
