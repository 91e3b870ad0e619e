"""Serving on the GPU: queries on several resident graphs, submitted together or as a
timed stream, are served with continuous batching; every result must equal the
single-query result (BFS depths and SSSP distances bit-exact, personalized PageRank
bitwise equal to a one-source mg_ppr_batch call), and later queries must join
batches already running."""
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graphs(mg):
    hosts = [mg.rmat_graph(12, 16, seed=21, weight_max=64), mg.grid_graph(48, seed=22, weight_max=16),
             mg.rmat_graph(10, 8, seed=23, weight_max=8)]
    dgs = [mg.Graph(h, symmetric=True) for h in hosts]
    yield hosts, dgs
    for d in dgs:
        d.close()


def _alone(mg, dg, kind, src):
    if kind == mg.MG_Q_BFS:
        return dg.bfs(src)
    if kind == mg.MG_Q_SSSP:
        return dg.sssp(src)
    return dg.ppr_batch([src])[0]


def _same(mg, kind, got, exp):
    if kind == mg.MG_Q_PPR:
        return np.array_equal(got.view(np.int32), exp.view(np.int32))
    return np.array_equal(got, exp)


def _queries(mg, hosts, count, seed):
    rng = np.random.default_rng(seed)
    qs = []
    for _ in range(count):
        gi = int(rng.integers(0, len(hosts)))
        kind = int(rng.choice([mg.MG_Q_BFS, mg.MG_Q_SSSP, mg.MG_Q_PPR], p=[0.6, 0.3, 0.1]))
        src = int(mg.pick_sources(hosts[gi], 1, seed=int(rng.integers(1 << 30)))[0])
        qs.append((gi, kind, src))
    return qs


@pytest.mark.parametrize("plain_kinds", [None, 0])
def test_serving_mixed_equals_alone(mg, device, graphs, plain_kinds):
    # default settings (SSSP as plain whole-GPU searches) and every kind batched
    hosts, dgs = graphs
    qs = _queries(mg, hosts, 150, seed=1)
    extra = {} if plain_kinds is None else {"plain_kinds": plain_kinds}
    with mg.Server(max_slots=16, steps_per_round=2, **extra) as srv:
        ids = [srv.add_graph(d) for d in dgs]
        outs, qids = [], []
        for gi, kind, src in qs:
            out = mg.PinnedArray(hosts[gi].n, np.float32 if kind == mg.MG_Q_PPR else np.int32)
            outs.append(out)
            qids.append(srv.submit(ids[gi], kind, src, out))
        infos = [srv.wait(q, 60000) for q in qids]
        st = srv.stats()
    assert st["completed"] == len(qs)
    for (gi, kind, src), out, info in zip(qs, outs, infos):
        exp = _alone(mg, dgs[gi], kind, src)
        assert _same(mg, kind, out.array, exp), (gi, kind, src)
        assert info["status"] == 0 and info["t_submit_ms"] <= info["t_admit_ms"] <= info["t_done_ms"]
        if kind != mg.MG_Q_PPR:
            reached = exp >= 0 if kind == mg.MG_Q_BFS else exp != 2**31 - 1
            assert info["vertices_reached"] == int(reached.sum())
            assert info["edges_traversed"] == int(np.diff(hosts[gi].row_offsets)[reached].sum())
    # continuous batching: with 16 slots per engine, some query was admitted while
    # an earlier query of the same engine was still running
    joined = 0
    for a, (qa, ia) in enumerate(zip(qs, infos)):
        for qb, ib in zip(qs[a + 1:], infos[a + 1:]):
            if qa[:2] == qb[:2] and ia["t_admit_ms"] < ib["t_admit_ms"] < ia["t_done_ms"]:
                joined += 1
                break
    assert joined > 0


@pytest.mark.parametrize("delta", [0, 1, 9, -1])
def test_serving_thin_searches_equal_alone(mg, device, graphs, delta):
    """thin_levels=1 serves every BFS / SSSP query as a thin search (one workgroup runs a
    whole query inside one launch, with delta-stepping buckets of width `delta` for
    SSSP, -1 unbucketed): results equal the plain searches bit for bit, with the same
    TEPS counts, and queries join while others of the same engine run."""
    hosts, dgs = graphs
    qs = [q for q in _queries(mg, hosts, 160, seed=5 + delta) if q[1] != mg.MG_Q_PPR]
    with mg.Server(thin_levels=1, thin_lanes=8, sssp_delta=delta) as srv:
        ids = [srv.add_graph(d) for d in dgs]
        assert srv.stats()["thin_engines"] == 2 * len(dgs)
        outs, qids = [], []
        for gi, kind, src in qs:
            outs.append(mg.PinnedArray(hosts[gi].n))
            qids.append(srv.submit(ids[gi], kind, src, outs[-1]))
        infos = [srv.wait(q, 60000) for q in qids]
    for (gi, kind, src), out, info in zip(qs, outs, infos):
        exp = _alone(mg, dgs[gi], kind, src)
        assert _same(mg, kind, out.array, exp), (gi, kind, src, delta)
        reached = exp >= 0 if kind == mg.MG_Q_BFS else exp != 2**31 - 1
        assert info["vertices_reached"] == int(reached.sum())
        assert info["edges_traversed"] == int(np.diff(hosts[gi].row_offsets)[reached].sum())
    joined = sum(1 for a, (qa, ia) in enumerate(zip(qs, infos)) for qb, ib in zip(qs[a + 1:], infos[a + 1:])
                 if qa[:2] == qb[:2] and ia["t_admit_ms"] < ib["t_admit_ms"] < ia["t_done_ms"])
    assert joined > 0


@pytest.mark.parametrize("slice_us", [0, 1, -1])
def test_serving_high_diameter_graph_runs_thin(mg, oracle, device, slice_us):
    """A 300 x 300 grid (about 600 BFS levels) is detected by the probe search and its
    BFS / SSSP engines run thin searches by default; the R-MAT graph beside it does
    not.  Results equal the oracle's with the default time slice, with a 1 us slice (a
    search pauses after every level and resumes in the next launch) and unsliced."""
    grid = mg.grid_graph(300, seed=41, weight_max=16)
    rm = mg.rmat_graph(12, 16, seed=42, weight_max=64)
    dgs = [mg.Graph(grid, symmetric=True), mg.Graph(rm, symmetric=True)]
    try:
        with mg.Server(thin_slice_us=slice_us) as srv:
            ids = [srv.add_graph(d) for d in dgs]
            assert srv.stats()["thin_engines"] == 2
            srcs = [int(s) for s in mg.pick_sources(grid, 12, seed=3)]
            outs, qids = [], []
            for i, s in enumerate(srcs):
                kind = mg.MG_Q_BFS if i % 2 == 0 else mg.MG_Q_SSSP
                outs.append(mg.PinnedArray(grid.n))
                qids.append(srv.submit(ids[0], kind, s, outs[-1]))
            for q in qids:
                srv.wait(q, 60000)
        for i, s in enumerate(srcs):
            exp = oracle.bfs(grid.row_offsets, grid.col_idx, s) if i % 2 == 0 else \
                oracle.sssp(grid.row_offsets, grid.col_idx, grid.values, s)
            assert np.array_equal(outs[i].array, exp), (i, s)
    finally:
        for d in dgs:
            d.close()


def test_serving_timed_stream(mg, device, graphs):
    hosts, dgs = graphs
    qs = _queries(mg, hosts, 200, seed=2)
    rng = np.random.default_rng(3)
    gaps = rng.exponential(1.0 / 400.0, len(qs))  # Poisson arrivals, 400 queries/s
    with mg.Server() as srv:
        ids = [srv.add_graph(d) for d in dgs]
        qids = [None] * len(qs)
        outs = [mg.PinnedArray(hosts[gi].n, np.float32 if k == mg.MG_Q_PPR else np.int32) for gi, k, _ in qs]

        def client():
            t = time.perf_counter()
            for i, (gi, kind, src) in enumerate(qs):
                t += gaps[i]
                d = t - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                qids[i] = srv.submit(ids[gi], kind, src, outs[i])

        th = threading.Thread(target=client)
        th.start()
        th.join()
        infos = [srv.wait(q, 60000) for q in qids]
    lat = np.array([i["t_done_ms"] - i["t_submit_ms"] for i in infos])
    assert np.all(lat > 0)
    for i in range(0, len(qs), 10):
        gi, kind, src = qs[i]
        assert _same(mg, kind, outs[i].array, _alone(mg, dgs[gi], kind, src)), qs[i]


def test_serving_small_hbm_budget(mg, device):
    hosts = [mg.rmat_graph(10, 8, seed=31, weight_max=16), mg.rmat_graph(10, 8, seed=32, weight_max=16)]
    dgs = [mg.Graph(h, symmetric=True) for h in hosts]
    budget = 3 << 20
    try:
        with mg.Server(hbm_budget=budget) as srv:
            ids = [srv.add_graph(d) for d in dgs]
            qs = _queries(mg, hosts, 60, seed=4)
            outs = [mg.PinnedArray(hosts[gi].n, np.float32 if k == mg.MG_Q_PPR else np.int32) for gi, k, _ in qs]
            qids = [srv.submit(ids[gi], k, s, o) for (gi, k, s), o in zip(qs, outs)]
            for q in qids:
                srv.wait(q, 60000)
            assert srv.stats()["device_bytes"] <= budget
        for (gi, kind, src), o in zip(qs, outs):
            assert _same(mg, kind, o.array, _alone(mg, dgs[gi], kind, src))
    finally:
        for d in dgs:
            d.close()


def test_serving_rejects_bad_queries(mg, device, graphs):
    hosts, dgs = graphs
    unweighted = mg.Graph(mg.rmat_graph(8, 4, seed=1), symmetric=True)
    try:
        with mg.Server() as srv:
            gid = srv.add_graph(dgs[0])
            uid = srv.add_graph(unweighted)
            for args in [(gid + 7, mg.MG_Q_BFS, 0), (gid, 5, 0), (gid, mg.MG_Q_BFS, hosts[0].n), (uid, mg.MG_Q_SSSP, 0)]:
                with pytest.raises(mg.MgError):
                    srv.submit(*args)
            with pytest.raises(mg.MgError):
                srv.wait(12345, 10)
            q = srv.submit(uid, mg.MG_Q_BFS, 0)  # no output buffer: result discarded
            assert srv.wait(q, 60000)["status"] == 0
    finally:
        unweighted.close()


@pytest.mark.parametrize("thin", [False, True])
def test_serving_output_memory_kinds(mg, device, graphs, thin):
    """A query's row reaches pinned host memory (mg_host_alloc), ordinary pageable host
    memory and device memory alike, for batched and thin engines and for personalized
    PageRank: every copy equals the single-query result."""
    import ctypes as C

    hosts, dgs = graphs
    hip = C.CDLL("libamdhip64.so")
    kinds = [mg.MG_Q_BFS, mg.MG_Q_SSSP] + ([] if thin else [mg.MG_Q_PPR])
    cfg = {"thin_levels": 1} if thin else {"thin_levels": -1}
    dev = []
    try:
        with mg.Server(**cfg) as srv:
            ids = [srv.add_graph(d) for d in dgs]
            runs = []
            for gi, d in enumerate(dgs):
                n = hosts[gi].n
                for kind in kinds:
                    src = int(mg.pick_sources(hosts[gi], 1, seed=gi * 7 + kind)[0])
                    dt = np.float32 if kind == mg.MG_Q_PPR else np.int32
                    pinned = mg.PinnedArray(n, dt)
                    pageable = np.full(n, -7, dt)
                    p = C.c_void_p()
                    assert hip.hipMalloc(C.byref(p), C.c_size_t(4 * n)) == 0
                    dev.append(p)
                    qids = [srv.submit(ids[gi], kind, src, pinned)]
                    for ptr, on_dev in ((pageable.ctypes.data, 0), (p.value, 1)):
                        q = C.c_int64()
                        mg.check(mg.lib().mg_server_submit(srv._h, ids[gi], kind, src, C.c_void_p(ptr), on_dev,
                                                           C.byref(q)))
                        qids.append(q.value)
                    runs.append((gi, kind, src, pinned, pageable, p, qids))
            for *_, qids in runs:
                for q in qids:
                    assert srv.wait(q, 60000)["status"] == 0
        for gi, kind, src, pinned, pageable, p, _ in runs:
            exp = _alone(mg, dgs[gi], kind, src)
            ondev = np.empty_like(pageable)
            assert hip.hipMemcpy(ondev.ctypes.data_as(C.c_void_p), p, C.c_size_t(ondev.nbytes), 2) == 0
            for got in (pinned.array, pageable, ondev):
                assert _same(mg, kind, got, exp), (gi, kind, src, thin)
    finally:
        for p in dev:
            hip.hipFree(p)


def test_serving_ppr_cohorts(mg, device, graphs):
    """Personalized PageRank queries start in cohorts: with ppr_batch_min = 4 and a long
    gate, three queries wait until a fourth arrives; then all four run together, and
    every row equals the single-query result bit for bit."""
    hosts, dgs = graphs
    srcs = [int(s) for s in mg.pick_sources(hosts[0], 4, seed=12)]
    with mg.Server(ppr_batch_min=4, ppr_gate_us=30_000_000) as srv:
        gid = srv.add_graph(dgs[0])
        outs = [mg.PinnedArray(hosts[0].n, np.float32) for _ in srcs]
        qids = [srv.submit(gid, mg.MG_Q_PPR, s, o) for s, o in zip(srcs[:3], outs[:3])]
        with pytest.raises(mg.MgError) as e:
            srv.wait(qids[0], 300)  # not started: the cohort is not full and the gate is 30 s
        assert e.value.code == mg.MG_ERR_TIMEOUT
        t4 = time.perf_counter()
        qids.append(srv.submit(gid, mg.MG_Q_PPR, srcs[3], outs[3]))
        infos = [srv.wait(q, 60000) for q in qids]
        assert time.perf_counter() - t4 < 20
    admitted = [i["t_admit_ms"] for i in infos]
    assert min(admitted) >= infos[3]["t_submit_ms"]  # nobody started before the fourth query
    for s, o in zip(srcs, outs):
        assert _same(mg, mg.MG_Q_PPR, o.array, _alone(mg, dgs[0], mg.MG_Q_PPR, s))


def test_serving_on_degree_ordered_twins(mg, device):
    """On graphs that keep a degree-ordered twin, the batched engines search the twin:
    sources are renamed on admission and rows mapped back on delivery, into pinned,
    pageable and device memory.  BFS / SSSP rows equal the plain graph's; personalized
    PageRank rows equal mg_ppr_batch on the same graph (which runs on the twin too),
    bit for bit."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    hosts = [mg.rmat_graph(12, 16, seed=51, weight_max=64), mg.rmat_graph(11, 8, seed=52, weight_max=16)]
    dgs = [mg.Graph(h, symmetric=True, relabel=True) for h in hosts]
    plain = [mg.Graph(h, symmetric=True) for h in hosts]
    dev = []
    try:
        with mg.Server(thin_levels=-1) as srv:
            ids = [srv.add_graph(d) for d in dgs]
            runs = []
            for i, (gi, kind, src) in enumerate(_queries(mg, hosts, 48, seed=53)):
                n = hosts[gi].n
                dt = np.float32 if kind == mg.MG_Q_PPR else np.int32
                how = i % 3
                if how == 0:
                    buf = mg.PinnedArray(n, dt)
                    q = srv.submit(ids[gi], kind, src, buf)
                else:
                    if how == 1:
                        buf = np.full(n, -7, dt)
                        ptr, on_dev = buf.ctypes.data, 0
                    else:
                        p = C.c_void_p()
                        assert hip.hipMalloc(C.byref(p), C.c_size_t(4 * n)) == 0
                        dev.append(p)
                        buf, ptr, on_dev = p, p.value, 1
                    qq = C.c_int64()
                    mg.check(mg.lib().mg_server_submit(srv._h, ids[gi], kind, src, C.c_void_p(ptr), on_dev,
                                                       C.byref(qq)))
                    q = qq.value
                runs.append((gi, kind, src, how, buf, dt, q))
            for *_, q in runs:
                assert srv.wait(q, 60000)["status"] == 0
        for gi, kind, src, how, buf, dt, _ in runs:
            if how == 0:
                got = buf.array
            elif how == 1:
                got = buf
            else:
                got = np.empty(hosts[gi].n, dt)
                assert hip.hipMemcpy(got.ctypes.data_as(C.c_void_p), buf, C.c_size_t(got.nbytes), 2) == 0
            assert _same(mg, kind, got, _alone(mg, dgs[gi], kind, src)), (gi, kind, src, how)
            if kind != mg.MG_Q_PPR:
                assert np.array_equal(got, _alone(mg, plain[gi], kind, src)), (gi, kind, src, how)
    finally:
        for p in dev:
            hip.hipFree(p)
        for d in dgs + plain:
            d.close()


def test_serving_plain_engines(mg, device, graphs):
    """plain_kinds: BFS / SSSP queries run one at a time as whole-GPU searches on a worker
    thread per (graph, kind) (PageRank stays batched); rows into pinned and pageable
    memory equal the single-query results, with the same TEPS counts."""
    import ctypes as C

    hosts, dgs = graphs
    qs = _queries(mg, hosts, 90, seed=61)
    direct = {}

    def direct_calls():  # the caller's own searches on the same graphs, beside the server's
        for gi, d in enumerate(dgs):
            s = int(mg.pick_sources(hosts[gi], 1, seed=70 + gi)[0])
            direct[gi] = (s, d.sssp(s), d.bfs(s))

    with mg.Server(plain_kinds=(1 << mg.MG_Q_BFS) | (1 << mg.MG_Q_SSSP)) as srv:
        ids = [srv.add_graph(d) for d in dgs]
        th = threading.Thread(target=direct_calls)
        th.start()
        outs, qids = [], []
        for i, (gi, kind, src) in enumerate(qs):
            dt = np.float32 if kind == mg.MG_Q_PPR else np.int32
            if i % 2:
                out = np.full(hosts[gi].n, -7, dt)
                q = C.c_int64()
                mg.check(mg.lib().mg_server_submit(srv._h, ids[gi], kind, src, C.c_void_p(out.ctypes.data), 0,
                                                   C.byref(q)))
                qids.append(q.value)
            else:
                out = mg.PinnedArray(hosts[gi].n, dt)
                qids.append(srv.submit(ids[gi], kind, src, out))
            outs.append(out)
        infos = [srv.wait(q, 60000) for q in qids]
        th.join()
        assert srv.stats()["completed"] == len(qs)
    for gi, (s, dist, depth) in direct.items():
        assert np.array_equal(dist, dgs[gi].sssp(s)) and np.array_equal(depth, dgs[gi].bfs(s))
    for (gi, kind, src), out, info in zip(qs, outs, infos):
        got = out if isinstance(out, np.ndarray) else out.array
        exp = _alone(mg, dgs[gi], kind, src)
        assert _same(mg, kind, got, exp), (gi, kind, src)
        if kind != mg.MG_Q_PPR:
            reached = exp >= 0 if kind == mg.MG_Q_BFS else exp != 2**31 - 1
            assert info["vertices_reached"] == int(reached.sum())
            assert info["edges_traversed"] == int(np.diff(hosts[gi].row_offsets)[reached].sum())
