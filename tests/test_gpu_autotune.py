"""mg_autotune: the chosen options must give the same results as the defaults (the
oracle's), the report must be consistent, and bad arguments are status codes.  The
tuned-options table is applied by graph class (CPU test)."""
import numpy as np
import pytest

INT_MAX = 2**31 - 1


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["rmat", "grid"])
def test_autotune_keeps_results(mg, oracle, device, kind):
    g = mg.rmat_graph(14, 16, seed=3, weight_max=64) if kind == "rmat" else mg.grid_graph(96, seed=3, weight_max=16)
    srcs = mg.pick_sources(g, 4, seed=1)
    with mg.Graph(g, symmetric=True) as dg:
        for what in ("bfs", "sssp", "bfs_batch", "sssp_batch"):
            best, rep = dg.autotune(what, srcs, reps=2)
            assert rep["candidates"] >= 5, (what, rep)
            assert 0 < rep["best_ms"] <= rep["start_ms"] * 1.0001, (what, rep)
            for j, s in enumerate(srcs):
                s = int(s)
                if what in ("bfs", "bfs_batch"):
                    exp = oracle.bfs(g.row_offsets, g.col_idx, s)
                    got = dg.bfs(s, best) if what == "bfs" else dg.bfs_batch(srcs, best)[j]
                else:
                    exp = oracle.sssp(g.row_offsets, g.col_idx, g.values, s)
                    got = dg.sssp(s, best) if what == "sssp" else dg.sssp_batch(srcs, best)[j]
                assert np.array_equal(got, exp), (what, s)


@pytest.mark.gpu
def test_autotune_rejects_bad_arguments(mg, device):
    g = mg.rmat_graph(10, 8, seed=1, weight_max=8)
    with mg.Graph(g, symmetric=True) as dg:
        for what, srcs in [("bfs", []), ("bfs", [g.n]), ("sssp", np.zeros(65, np.int32))]:
            with pytest.raises(mg.MgError) as e:
                dg.autotune(what, srcs)
            assert e.value.code == mg.MG_ERR_INVALID_ARG
        with pytest.raises(mg.MgError):
            dg.autotune("bfs", [0], reps=0)


def test_tuned_table_is_applied_by_graph_class(mg):
    from mirrorofmapgraph_amd import tuning

    g = mg.rmat_graph(12, 16, seed=1, weight_max=64)
    key = tuning.graph_class(g)
    table = {key: {"sssp": {"options": {"sssp_delta": 3, "level_grid": 1024, "not_a_tunable": 7}}}}
    o, applied = tuning.tuned_options(g, "sssp", table=table)
    assert applied == {"sssp_delta": 3, "level_grid": 1024}
    assert o.sssp_delta == 3 and o.level_grid == 1024 and o.alpha == 14
    o2, applied2 = tuning.tuned_options(g, "bfs", table=table)
    assert applied2 == {} and o2.sssp_delta == 0
    grid = mg.grid_graph(64, seed=1, weight_max=16)
    assert tuning.graph_class(grid) != key and "flat" in tuning.graph_class(grid) and "hub" in key
    # the committed table parses and names only tunable fields
    for cls, kinds in tuning.load_table().items():
        for what, entry in kinds.items():
            assert what in mg.TUNE_KINDS and set(entry["options"]) <= set(mg.TUNABLE), (cls, what)
