"""The command-line driver (reference bfs.cu interface) on the reference's fixture
graphs: its output file must equal the oracle's labels for the same source."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import load_mtx

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
DRIVER = ROOT / "mirrorofmapgraph_amd" / "mapgraph_run"
GOLDEN = ROOT / "tests" / "golden"


def _run(args, tmp_path):
    out = tmp_path / "out.txt"
    r = subprocess.run([str(DRIVER), *args, "-o", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout, np.loadtxt(out, dtype=np.int64, ndmin=1)


@pytest.mark.parametrize("name", ["small.mtx", "chesapeake.mtx"])
def test_driver_bfs_sssp_cc(mg, oracle, device, tmp_path, name):
    g = load_mtx(mg, name)
    stdout, d = _run(["bfs", "-g", str(GOLDEN / name), "-p", "src=1"], tmp_path)  # origin 1: vertex 0
    exp = oracle.bfs(g.row_offsets, g.col_idx, 0)
    assert np.array_equal(d, exp)
    assert f"Traversed edge: {oracle.teps_edges(g.row_offsets, exp)}" in stdout
    _, s = _run(["sssp", "-g", str(GOLDEN / name), "-p", "src=1"], tmp_path)
    assert np.array_equal(s, oracle.sssp(g.row_offsets, g.col_idx, g.values, 0))
    _, c = _run(["cc", "-g", str(GOLDEN / name)], tmp_path)
    assert np.array_equal(c, oracle.cc(g.row_offsets, g.col_idx))


def test_driver_random_sources_batch(mg, oracle, device, tmp_path):
    stdout, d = _run(["bfs", "-g", str(GOLDEN / "chesapeake.mtx"), "-s", "RANDOM", "-p", "num_src=300,seed=5"],
                     tmp_path)
    assert "batch of 256 sources" in stdout and "batch of 44 sources" in stdout
    g = load_mtx(mg, "chesapeake.mtx")
    last = int(mg.pick_sources(g, 300, seed=5)[-1])
    assert np.array_equal(d, oracle.bfs(g.row_offsets, g.col_idx, last))


def test_driver_verbose_level_trace(mg, oracle, device, tmp_path):
    """verbose=1 prints one line per level (the reference's per-iteration frontier
    sizes); for BFS the next frontier of level L is the set of vertices at depth L+1."""
    out = tmp_path / "out.txt"
    r = subprocess.run([str(DRIVER), "bfs", "-g", str(GOLDEN / "chesapeake.mtx"), "-p", "src=1,verbose=1", "-o",
                        str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [ln for ln in r.stderr.splitlines() if ln.startswith("bfs level ")]
    g = load_mtx(mg, "chesapeake.mtx")
    depth = oracle.bfs(g.row_offsets, g.col_idx, 0)
    levels = int(depth.max()) + 1  # the last level finds nothing
    assert len(lines) == levels
    for L, ln in enumerate(lines):
        nf = int(ln.split("next frontier ")[1].split()[0])
        assert nf == int((depth == L + 1).sum()), ln
