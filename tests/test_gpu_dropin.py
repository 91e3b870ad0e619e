"""The C++ drop-in API (include/mapgraph/csr_problem.h, enactor.h, programs/*.h) as a
MapGraph driver uses it: tests/dropin/mapgraph_dropin (built by the package build)
runs CsrProblem::FromHostProblem -> EnactorVertexCentric::EnactIterativeSearch ->
CsrProblem::ExtractResults (the sequence of Algorithms/BFS/bfs.cu:428-455) for the
four ported programs, and its output is compared with the CPU oracle: BFS depths,
SSSP distances and CC labels bit-exact, PageRank (the reference program) within
1e-6 in normalised units."""
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_mtx

BIN = ROOT / "tests" / "dropin" / "mapgraph_dropin"


def _write_graph(path, g):
    with open(path, "wb") as f:
        np.array([g.n, g.m, 0 if g.values is None else 1], np.int64).tofile(f)
        g.row_offsets.astype(np.int32).tofile(f)
        g.col_idx.astype(np.int32).tofile(f)
        if g.values is not None:
            g.values.astype(np.int32).tofile(f)


def _dropin(tmp_path, prog, g, params="", dtype=np.int32):
    gp, op = tmp_path / "graph.bin", tmp_path / "out.bin"
    _write_graph(gp, g)
    cmd = [str(BIN), prog, str(gp), str(op)] + (["-p", params] if params else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Kernel time took" in r.stdout and "Total iteration" in r.stdout
    out = np.fromfile(op, dtype=dtype)
    assert out.size == g.n
    return out


def test_dropin_driver_is_built():
    """The build compiles the drop-in driver; without arguments it prints its usage
    and exits 1 (no device is touched)."""
    assert BIN.exists(), f"{BIN} is not built; run the package build"
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["small.mtx", "trivialGraph.mtx", "chesapeake.mtx"])
def test_dropin_fixtures(mg, oracle, device, tmp_path, name):
    g = load_mtx(mg, name)
    for directed in (1, 0):
        for src in range(min(g.n, 3)):
            p = f"src={src + 1},origin=1,directed={directed}"  # 1-based sources, as the reference drivers
            d = _dropin(tmp_path, "bfs", g, p)
            assert np.array_equal(d, oracle.bfs(g.row_offsets, g.col_idx, src, directed=directed)), (directed, src)
            s = _dropin(tmp_path, "sssp", g, p)
            exp = oracle.sssp(g.row_offsets, g.col_idx, g.values, src, directed=directed)
            assert np.array_equal(s, exp), (directed, src)
    # cc.cu:289 runs CC with directed = false
    assert np.array_equal(_dropin(tmp_path, "cc", g, "directed=0"), oracle.cc(g.row_offsets, g.col_idx))


@pytest.mark.gpu
def test_dropin_rmat16(mg, oracle, device, tmp_path):
    g = mg.rmat_graph(16, 16, seed=42, weight_max=64)
    assert np.array_equal(_dropin(tmp_path, "bfs", g, "src=1"), oracle.bfs(g.row_offsets, g.col_idx, 0))
    assert np.array_equal(_dropin(tmp_path, "sssp", g, "src=1"),
                          oracle.sssp(g.row_offsets, g.col_idx, g.values, 0))
    assert np.array_equal(_dropin(tmp_path, "cc", g, "directed=0"), oracle.cc(g.row_offsets, g.col_idx))
    # iter_num caps the levels, as in the reference loop (enactor:2526)
    full = oracle.bfs(g.row_offsets, g.col_idx, 0)
    got = _dropin(tmp_path, "bfs", g, "src=1,iter_num=2")
    assert np.array_equal(got, np.where(full <= 2, full, -1))


@pytest.mark.gpu
def test_dropin_pagerank_program(mg, oracle, device, tmp_path):
    g = mg.rmat_graph(13, 16, seed=9, symmetric=False)
    got = _dropin(tmp_path, "pagerank", g, "", dtype=np.float32)
    exp, _ = oracle.pr_gas(g.row_offsets, g.col_idx)
    assert float(np.max(np.abs(got.astype(np.float64) - exp))) / g.n <= 1e-6
    # the activation known answer of the oracle's pinning test, through the drop-in
    edges = [(0, 1), (1, 2), (2, 3)] + [(1, s) for s in range(4, 16)]
    small = mg.csr_from_edges(16, np.array([e[0] for e in edges], np.int32), np.array([e[1] for e in edges], np.int32),
                              mg.MG_CSR_SORT_COLS)
    r = _dropin(tmp_path, "pagerank", small, "", dtype=np.float32)
    y2 = 0.15 + 0.85 * 0.2775 / 13
    assert np.allclose(r, [0.15, 0.2775, y2, 0.2775] + [y2] * 12, rtol=0, atol=1e-6)
