import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = Path(__file__).resolve().parent / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def mg():
    import mirrorofmapgraph_amd as m

    m.lib()  # raises when the library has not been built
    return m


@pytest.fixture(scope="session")
def oracle():
    import oracle as o

    o.lib()
    return o


@pytest.fixture(scope="session")
def device(mg):
    """The GPU; GPU tests fail loudly (never fall back) when none is visible."""
    if mg.device_count() == 0:
        pytest.fail("no HIP device visible for a gpu-marked test")
    mg.check(mg.lib().mg_device_init(0))
    return 0


def load_mtx(mg, name):
    """A reference fixture loaded as the drivers do: BuildMarketGraph(file, csr, false)
    -> CsrGraph::FromCoo (no symmetrize, no dedupe, stable row grouping)."""
    n, s, d, v, _ = mg.read_mtx(GOLDEN / name)
    return mg.csr_from_edges(n, s, d, 0, vals=v)


def random_graph(mg, n, m, seed, symmetric=True, wmax=None, loops=False):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, n, m, dtype=np.int32)
    d = rng.integers(0, n, m, dtype=np.int32)
    flags = mg.MG_CSR_DEDUPE | mg.MG_CSR_SORT_COLS | (0 if loops else mg.MG_CSR_DROP_LOOPS)
    if symmetric:
        flags |= mg.MG_CSR_SYMMETRIZE
    g = mg.csr_from_edges(n, s, d, flags)
    if wmax:
        g.values = mg.hash_weights(g, seed + 1, wmax)
    return g
