#!/usr/bin/env python3
"""Benchmark: SSSP from vertex 0 + connected components on the seeded R-MAT scale-22
graph (edge factor 16, (0.57, 0.19, 0.19, 0.05), symmetrized, self-loops and
duplicates removed, int32 CSR, weights in [1, 64]) on one MI355X.

One step = one SSSP search from vertex 0 plus one CC labelling, graph resident in
HBM, results left on the device.  value = (traversed edges of SSSP + of CC) / wall
time of the K timed steps, in GTEPS (traversed edges: the reference's count, sum of
out-degrees of reached vertices, bfs.cu:483-493).  BFS from vertex 0 is timed
beside it.  The roofline object covers the SSSP search, the step's dominant part:
algorithmic bytes 12*m_r + 24*n_r per search over its device time (HIP events on
the graph's stream, around every kernel of the search).

After the timed steps every result is compared with the CPU oracle on the same
graph (as the reference drivers compare CPU and GPU labels in the same run,
bfs.cu:495-504, sssp.cu:433-441); a mismatch prints the first differing vertices
and exits with status 1.  The extra parts (other workloads of the benchmark suite)
are checked the same way.  Part timings are the median of --reps runs after two
warm-ups.  The CPU baseline runs the oracle's restatements of the reference CPU code
on one host core (the process is pinned to it while they run).

Prints one JSON line.  Only one GPU is supported.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
# Serving runs one stream per (graph, query kind) engine (18 for six graphs) plus three
# launch streams per thin engine (12 for the two grids); streams that share a hardware
# queue run in order, so short rounds would queue behind long launches.  Read at runtime
# initialisation, so it is set before the library loads.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "32")

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
INT_MAX = 2**31 - 1
PR_TOL = 1e-6  # PageRank: max-abs in normalised units p = r / n
SERVE_CONFIG = {}  # mg_server_config fields the serving part sets (defaults otherwise)
METRIC = "GTEPS & %HBM roofline, R-MAT s22 BFS/SSSP; batched vs plain; p99 query ms"


def bytes_bfs(m_r, n_r):
    return 8 * m_r + 20 * n_r


def bytes_sssp(m_r, n_r):
    return 12 * m_r + 24 * n_r


def bytes_cc(m, n):
    return 8 * m + 20 * n


COUNTERS = ROOT / "profiles" / "rmat22_sssp_cc_bfs_counters.md"


class Mismatch(Exception):
    pass


def require_equal(what: str, got: np.ndarray, exp: np.ndarray) -> None:
    bad = np.nonzero(got != exp)[0]
    if bad.size:
        raise Mismatch(f"{what}: GPU result differs from the CPU oracle at {bad.size} vertices, first "
                       f"{bad[:8].tolist()} (gpu {got[bad[:8]].tolist()}, oracle {exp[bad[:8]].tolist()})")


def median_ms(fn, reps: int, warmup: int = 2):
    """Runs fn() (returning (result, stats)) warmup + reps times; median device ms and the last stats."""
    for _ in range(warmup):
        fn()
    ms, st = [], None
    for _ in range(reps):
        _, st = fn()
        ms.append(st.device_ms)
    return float(np.median(ms)), st


def measured_traffic(kind: str):
    """HBM bytes of one search as measured by rocprofv3 FETCH_SIZE / WRITE_SIZE runs of
    `bench.py --pmc-pass` (scaled by a copy-kernel calibration); None when absent."""
    try:
        for line in COUNTERS.read_text().splitlines():
            if line.startswith("machine-readable: "):
                return int(json.loads(line.split(": ", 1)[1])[kind]["traffic_bytes"])
    except (OSError, KeyError, ValueError):
        pass
    return None


def cpu_info() -> dict:
    """The host CPU as /proc/cpuinfo describes it (first processor) and the cores this process may use."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if not line.strip():
                break
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("vendor_id", "model name", "cpu family", "model", "stepping", "cpu MHz", "cache size"):
                info[k] = v.strip()
    except OSError:
        info["model name"] = platform.processor() or "unknown"
    info["logical_cpus"] = os.cpu_count()
    try:
        info["usable_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    return info


class pinned_core:
    """Pins the calling thread to one core (as `taskset -c <core>` would) while the block runs."""

    def __enter__(self):
        try:
            self.saved = os.sched_getaffinity(0)
            self.core = max(self.saved)
            os.sched_setaffinity(0, {self.core})
        except (AttributeError, OSError):
            self.saved, self.core = None, None
        return self

    def __exit__(self, *exc):
        if self.saved is not None:
            os.sched_setaffinity(0, self.saved)


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 5) -> float:
    """Device-to-device copy of nbytes (a read and a write of every byte) timed with HIP
    events; best of reps, in GB/s of read + write traffic: the achievable HBM rate."""
    hip = C.CDLL("libamdhip64.so")  # the runtime the engine already loaded
    p, q, e0, e1 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(q), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed for the copy-bandwidth probe")
    hip.hipEventCreate(C.byref(e0))
    hip.hipEventCreate(C.byref(e1))
    best = None
    try:
        for _ in range(reps + 1):
            hip.hipEventRecord(e0, None)
            if hip.hipMemcpyAsync(q, p, C.c_size_t(nbytes), 3, None):  # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpyAsync failed")
            hip.hipEventRecord(e1, None)
            hip.hipEventSynchronize(e1)
            ms = C.c_float()
            hip.hipEventElapsedTime(C.byref(ms), e0, e1)
            best = ms.value if best is None else min(best, ms.value)
    finally:
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        hip.hipFree(p)
        hip.hipFree(q)
    return 2 * nbytes / (best * 1e-3) / 1e9


def rmat16_part(mg, oracle, seed: int, reps: int, log) -> dict:
    """BFS from vertex 0 on the scale-16 R-MAT (symmetrized), depths checked against the oracle."""
    host = mg.rmat_graph(16, 16, seed=seed)
    with mg.Graph(host, symmetric=True) as dg:
        t, st = median_ms(lambda: dg.bfs(0, stats=True, copy=False), reps)
        require_equal("rmat16 BFS from 0", dg.bfs(0), oracle.bfs(host.row_offsets, host.col_idx, 0))
    b = bytes_bfs(st.edges_traversed, st.vertices_reached)
    log(f"rmat16 bfs: {t:.3f} ms")
    return {"n": host.n, "m": host.m, "device_ms": t, "levels": st.iterations, "gteps": st.edges_traversed / t / 1e6,
            "hbm_frac": b / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, "checked": True}


def grid_part(mg, oracle, seed: int, reps: int, log) -> dict:
    """BFS and SSSP from vertex 0 (a corner) of the 2048x2048 4-neighbour grid with 5% of the edges deleted
    (seeded) and weights 1..16: about 4,100 levels, the high-diameter case.  Results checked against the oracle."""
    from mirrorofmapgraph_amd.tuning import tuned_options

    host = mg.grid_graph(2048, seed=seed, weight_max=16)
    out = {"n": host.n, "m": host.m}
    with mg.Graph(host, symmetric=True) as dg:
        for kind, fn, nb in [("bfs", dg.bfs, bytes_bfs), ("sssp", dg.sssp, bytes_sssp)]:
            o, tuned = tuned_options(host, kind)
            t, st = median_ms(lambda: fn(0, o, stats=True, copy=False), reps)
            b = nb(st.edges_traversed, st.vertices_reached)
            out[kind] = {"device_ms": t, "levels": st.iterations, "gteps": st.edges_traversed / t / 1e6,
                         "edges_examined": st.edges_examined, "us_per_level": 1e3 * t / max(st.iterations, 1),
                         "hbm_frac": b / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, "tuned_options": tuned}
            log(f"grid2048 {kind}: {t:.2f} ms, {st.iterations} levels")
        require_equal("grid2048 BFS from 0", dg.bfs(0, tuned_options(host, "bfs")[0]),
                      oracle.bfs(host.row_offsets, host.col_idx, 0))
        require_equal("grid2048 SSSP from 0", dg.sssp(0, tuned_options(host, "sssp")[0]),
                      oracle.sssp(host.row_offsets, host.col_idx, host.values, 0))
    out["checked"] = True
    return out


def calibration_copy(nbytes: int) -> None:
    """A device-to-device copy of nbytes (the runtime's copy kernel): the known traffic that scales the
    FETCH_SIZE / WRITE_SIZE counters of the same run."""
    hip = C.CDLL("libamdhip64.so")  # the runtime the engine already loaded
    p, q = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(q), C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(q, p, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(p)
    hip.hipFree(q)


def pagerank_part(mg, oracle, seed: int, reps: int, log, relabel: bool = True) -> dict:
    """PageRank on the directed scale-22 R-MAT (dangling vertices kept), all-vertex Jacobi, damping 0.85,
    FP32, 50 iterations (the L1 stop at 1e-6 never triggers in FP32 at this size).  Ranks are compared with
    the FP32 Jacobi oracle (normalised max-abs <= 1e-6) and their error against the FP64 oracle is stated."""
    host = mg.rmat_graph(22, 16, seed=seed, symmetric=False)
    with mg.Graph(host, build_csc=True, relabel=relabel) as dg:
        o = mg.options(pr_mode=mg.MG_PR_ALL, iter_num=50, pr_tol=1e-6)
        t, st = median_ms(lambda: dg.pagerank(o, stats=True, copy=False), reps)
        got = dg.pagerank(o)
    exp, it, _ = oracle.pr_jacobi_f32(host.row_offsets, host.col_idx, iters=50, tol=1e-6)
    ref64, _, _ = oracle.pr_jacobi_f64(host.row_offsets, host.col_idx, iters=50)
    err32 = float(np.max(np.abs(got.astype(np.float64) - exp))) / host.n
    err64 = float(np.max(np.abs(got.astype(np.float64) - ref64))) / host.n
    if it != st.iterations or err32 > PR_TOL or err64 > PR_TOL:
        raise Mismatch(f"pagerank s22d: {st.iterations} iterations (oracle {it}), normalised max-abs error "
                       f"{err32:.3g} vs the FP32 oracle, {err64:.3g} vs FP64 (bound {PR_TOL})")
    per_it = t / st.iterations
    b = 8 * host.m + 16 * host.n
    log(f"pagerank s22d: {t:.2f} ms for {st.iterations} iterations, err32 {err32:.2e} err64 {err64:.2e}")
    return {"n": host.n, "m": host.m, "iterations": st.iterations, "device_ms": t, "ms_per_iteration": per_it,
            "gteps": host.m * st.iterations / t / 1e6, "algorithmic_bytes_per_iteration": b,
            "hbm_frac": b / (per_it * 1e-3) / 1e9 / HBM_PEAK_GBS, "max_abs_err_vs_fp32_oracle": err32,
            "max_abs_err_vs_fp64_oracle": err64, "error_units": "normalised p = r / n", "checked": True}


def batched_part(mg, oracle, seed: int, reps: int, log) -> dict:
    """256 BFS and 256 SSSP sources (seeded draws among non-isolated vertices) on the R-MAT scale-20 graph
    and the 2048x2048 grid: one batched call each against the same 256 sources run one by one (plain, every
    source timed).  Every batched row is compared with its plain search, and sampled rows with the oracle."""
    out = {}
    hosts = [("rmat20", mg.rmat_graph(20, 16, seed=seed, weight_max=64)),
             ("grid2048", mg.grid_graph(2048, seed=seed, weight_max=16))]
    sources = {name: mg.pick_sources(host, 256, seed=seed) for name, host in hosts}
    out["both_graphs_one_search"] = multi_graph_part(mg, hosts, sources, reps, log)
    from mirrorofmapgraph_amd.tuning import tuned_options

    for name, host in hosts:
        srcs = sources[name]
        # the R-MAT graph keeps its degree-ordered twin (plain and batched searches both run there)
        with mg.Graph(host, symmetric=True, relabel=name.startswith("rmat")) as dg:
            res = {"n": host.n, "m": host.m, "sources": 256, "relabel": name.startswith("rmat")}
            for kind, batch, plain, orc in [
                    ("bfs", dg.bfs_batch, dg.bfs, lambda s: oracle.bfs(host.row_offsets, host.col_idx, s)),
                    ("sssp", dg.sssp_batch, dg.sssp,
                     lambda s: oracle.sssp(host.row_offsets, host.col_idx, host.values, s))]:
                # each form with the autotuner's options for this graph class and search kind
                ob, tuned_b = tuned_options(host, kind + "_batch")
                op, tuned_p = tuned_options(host, kind)
                tb, st = median_ms(lambda: batch(srcs, ob, stats=True, copy=False), reps)
                te = st.edges_traversed
                rows = batch(srcs, ob)
                plain(int(srcs[0]), op, copy=False)
                tp = 0.0
                for j, v in enumerate(srcs):
                    row, pst = plain(int(v), op, stats=True)
                    tp += pst.device_ms
                    require_equal(f"{name} {kind} batched row {j} vs its plain search", rows[j], row)
                for j in (0, 255):
                    require_equal(f"{name} {kind} row {j} vs oracle", rows[j], orc(int(srcs[j])))
                del rows
                res[kind] = {"batched_ms": tb, "batched_gteps": te / tb / 1e6, "plain_ms": tp,
                             "plain_gteps": te / tp / 1e6, "speedup": tp / tb, "levels": st.iterations,
                             "plain_sources_timed": 256, "checked": True,
                             "tuned_options": {"batched": tuned_b, "plain": tuned_p}}
                log(f"batched {name} {kind}: {tb:.2f} ms vs plain {tp:.1f} ms ({tp / tb:.1f}x)")
        out[name] = res
    return out


def multi_graph_part(mg, hosts, sources, reps: int, log) -> dict:
    """Both batched graphs in one search: their disjoint union (mg_graph_union) with row j carrying the j-th
    source of each graph (256 + 256 sources in 256 rows), against the two separate batched calls; every
    graph's rows are compared with its separate call."""
    dgs = [mg.Graph(h, symmetric=True) for _, h in hosts]
    res = {"graphs": [name for name, _ in hosts], "sources_per_graph": 256}
    try:
        u = mg.Graph.union(dgs)
        srcs = [sources[name] for name, _ in hosts]
        with u:
            for kind in ("bfs", "sssp"):
                one = lambda: u.parts_batch(kind, srcs, stats=True, copy=False)
                t_one, st = median_ms(one, reps)
                parts = u.parts_batch(kind, srcs)
                t_sep = 0.0
                for i, dg in enumerate(dgs):
                    fn = dg.bfs_batch if kind == "bfs" else dg.sssp_batch
                    t, _ = median_ms(lambda: fn(srcs[i], stats=True, copy=False), reps)
                    t_sep += t
                    require_equal(f"{hosts[i][0]} {kind} rows of the one-search union vs the graph's own batch",
                                  parts[i].ravel(), fn(srcs[i]).ravel())
                del parts
                res[kind] = {"one_search_ms": t_one, "separate_ms": t_sep, "speedup": t_sep / t_one,
                             "gteps": st.edges_traversed / t_one / 1e6, "levels": st.iterations, "checked": True}
                log(f"union rmat20+grid2048 {kind}: {t_one:.1f} ms in one search vs {t_sep:.1f} ms separately")
    finally:
        for d in dgs:
            d.close()
    return res


def varied_sources_part(mg, oracle, dg, g, seed: int, log) -> dict:
    """BFS and SSSP on the scale-22 graph from 16 different sources in turn (each search differs from the
    previous one, unlike the repeated source of the main step); two of them checked against the oracle."""
    srcs = [int(s) for s in mg.pick_sources(g, 16, seed=seed + 1)]
    out = {"sources": len(srcs)}
    for kind, fn, nb in [("bfs", dg.bfs, bytes_bfs), ("sssp", dg.sssp, bytes_sssp)]:
        fn(srcs[-1], copy=False)
        ms, te, by = [], 0, 0.0
        for s in srcs:
            _, st = fn(s, stats=True, copy=False)
            ms.append(st.device_ms)
            te += st.edges_traversed
            by += nb(st.edges_traversed, st.vertices_reached)
        t = float(np.sum(ms))
        out[kind] = {"mean_device_ms": t / len(srcs), "gteps": te / t / 1e6,
                     "hbm_frac": by / (t * 1e-3) / 1e9 / HBM_PEAK_GBS}
        log(f"s22 {kind} over 16 sources: {t / len(srcs):.3f} ms mean")
    for s in srcs[:2]:
        require_equal(f"rmat22 BFS from {s}", dg.bfs(s), oracle.bfs(g.row_offsets, g.col_idx, s))
        require_equal(f"rmat22 SSSP from {s}", dg.sssp(s), oracle.sssp(g.row_offsets, g.col_idx, g.values, s))
    out["checked"] = True
    return out


def serving_part(mg, seed: int, seconds: float, rate: float, log, mix=(0.6, 0.3, 0.1), graphs=None,
                 cfg=None, relabel: bool = True) -> dict:
    """Poisson query stream (seeded) over six resident graphs -- R-MAT scales 16, 18, 20, 22 (symmetrized,
    weights 1..64; with the degree-ordered twin unless relabel is False) and 1024^2 / 2048^2 grids (weights
    1..16) -- with 60% BFS, 30% SSSP and 10% personalized PageRank (FP32, 20 iterations) from sources drawn
    among non-isolated vertices.  Latency = result copied to pinned host memory minus scheduled arrival.
    Every 25th query's result is compared with the same query run alone afterwards."""
    import threading

    hosts = [mg.rmat_graph(s, 16, seed=seed + s, weight_max=64) for s in (16, 18, 20, 22)]
    hosts += [mg.grid_graph(w, seed=seed + w, weight_max=16) for w in (1024, 2048)]
    names = ["rmat16", "rmat18", "rmat20", "rmat22", "grid1024", "grid2048"]
    if graphs:  # a subset of the six, for isolating engines
        keep = [names.index(x) for x in graphs]
        hosts, names = [hosts[i] for i in keep], [names[i] for i in keep]
    dgs = [mg.Graph(h, symmetric=True, relabel=relabel and x.startswith("rmat")) for h, x in zip(hosts, names)]
    rng = np.random.default_rng(seed)
    nq = int(rng.poisson(rate * seconds))
    gaps = rng.exponential(1.0 / rate, nq)
    gsel = rng.integers(0, len(hosts), nq)
    kinds = rng.choice([mg.MG_Q_BFS, mg.MG_Q_SSSP, mg.MG_Q_PPR], p=list(mix), size=nq)
    pools = [mg.pick_sources(h, 4096, seed=seed + i) for i, h in enumerate(hosts)]
    srcs = [int(pools[g][i % 4096]) for i, g in enumerate(gsel)]
    bufs = [[mg.PinnedArray(h.n) for _ in range(4)] for h in hosts]  # result buffers, reused round-robin
    checked = {i: mg.PinnedArray(hosts[gsel[i]].n) for i in range(0, nq, 25)}  # kept for the alone check
    qids = [0] * nq
    arrive = np.cumsum(gaps) * 1e3  # ms after the stream starts
    cfg = dict(SERVE_CONFIG, **(cfg or {}))
    with mg.Server(**cfg) as srv:
        ids = [srv.add_graph(d) for d in dgs]
        t_start = srv.now_ms()

        def client():
            t = time.perf_counter()
            for i in range(nq):
                t += gaps[i]
                d = t - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                g = int(gsel[i])
                qids[i] = srv.submit(ids[g], int(kinds[i]), srcs[i], checked.get(i, bufs[g][i % 4]))

        th = threading.Thread(target=client)
        th.start()
        th.join()
        infos = [srv.wait(q, 600000) for q in qids]
        st = srv.stats()
    for i, buf in checked.items():
        dg, kind, s = dgs[gsel[i]], int(kinds[i]), srcs[i]
        if kind == mg.MG_Q_PPR:
            alone = dg.ppr_batch([s])[0].view(np.int32)
        else:
            alone = dg.bfs(s) if kind == mg.MG_Q_BFS else dg.sssp(s)
        require_equal(f"serving query {i} ({names[gsel[i]]}, kind {kind}, source {s}) vs alone", buf.array, alone)
    for d in dgs:
        d.close()
    # latency from the scheduled arrival (a late submit call counts against the server)
    lat = np.array([i["t_done_ms"] - (t_start + arrive[k]) for k, i in enumerate(infos)])
    t_end = max(i["t_done_ms"] for i in infos)
    te = sum(i["edges_traversed"] for i in infos)
    # steady state: queries arriving after the first second and before the arrival window ends
    steady = (arrive > 1000.0) & (arrive <= seconds * 1e3)
    done_rel = np.array([i["t_done_ms"] - t_start for i in infos])
    span = (arrive[-1] - 1000.0) / 1e3 if nq else 0.0
    steady_rate = float(((done_rel > 1000.0) & (done_rel <= arrive[-1])).sum() / span) if span > 0 else None
    steady_in = float(((arrive > 1000.0) & (arrive <= arrive[-1])).sum() / span) if span > 0 else None
    res = {"queries": nq, "rate_qps": rate, "arrival_window_s": seconds, "mix_bfs_sssp_ppr": list(mix), "p50_ms": float(np.percentile(lat, 50)),
           "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
           "steady_p50_ms": float(np.percentile(lat[steady], 50)) if steady.any() else None,
           "steady_p99_ms": float(np.percentile(lat[steady], 99)) if steady.any() else None,
           "aggregate_gteps": te / (t_end - t_start) / 1e6, "completed_qps": nq / ((t_end - t_start) / 1e3),
           # the stream as it arrived, and how long the server took to drain it after the last arrival
           # (completed_qps counts that drain: a server that keeps up completes nq / (window + drain))
           "arrived_qps": nq / (arrive[-1] / 1e3) if nq else 0.0, "drain_ms": t_end - (t_start + arrive[-1]),
           # completions per second while the stream is still arriving (from its first second to its last
           # arrival): the service rate under the sustained load, equal to the arrival rate when the server keeps
           # up and below it when a backlog grows
           "steady_completed_qps": steady_rate,
           "steady_arrived_qps": steady_in,
           "device_bytes": st["device_bytes"], "graphs": names, "checked_queries": len(checked),
           "thin_engines": st["thin_engines"], "dispatcher_busy_frac": st["dispatch_busy_ms"] / (t_end - t_start),
           "dispatcher_passes": st["dispatch_passes"], "rounds": st["rounds"], "server_config": cfg,
           "rmat_twins": relabel}
    # service time: admitted into a slot / lane until done (the rest of the latency is waiting)
    svc = np.array([i["t_done_ms"] - i["t_admit_ms"] for i in infos])
    per = {}
    for k, kn in [(mg.MG_Q_BFS, "bfs"), (mg.MG_Q_SSSP, "sssp"), (mg.MG_Q_PPR, "ppr")]:
        for g, gn in enumerate(names):
            sel = (kinds == k) & (gsel == g)
            if sel.any():
                per[f"{kn}_{gn}"] = {"n": int(sel.sum()), "p50_ms": float(np.percentile(lat[sel], 50)),
                                     "p99_ms": float(np.percentile(lat[sel], 99)),
                                     "service_p50_ms": float(np.percentile(svc[sel], 50))}
    res["per_kind_graph"] = per
    log(f"serving: {nq} queries, p50 {res['p50_ms']:.2f} ms, p99 {res['p99_ms']:.2f} ms, "
        f"{res['aggregate_gteps']:.1f} GTEPS, {res['completed_qps']:.0f} q/s "
        f"(steady: {res['steady_completed_qps'] or 0:.0f} completed, {res['steady_arrived_qps'] or 0:.0f} arrived)")
    return res


def cpu_baseline(mg, oracle, g, src, te, log):
    """The reference CPU path restated (oracle): SSSP (binary-heap restatement of CPUSSSP: identical
    distances) + CPUCC on the same graph, one pinned host core.  Also the verbatim O(n^2) CPUSSSP on the
    scale-16 graph.  Returns the baseline record, the oracle's SSSP / CC answers (reused as the checker) and
    the scale-16 graph with its quadratic-CPUSSSP distances."""
    s16 = mg.rmat_graph(16, 16, seed=42, weight_max=64)
    with pinned_core() as pin:
        t = time.perf_counter()
        d_ref = oracle.sssp(g.row_offsets, g.col_idx, g.values, src)
        t_sssp = time.perf_counter() - t
        t = time.perf_counter()
        cc_ref = oracle.cc(g.row_offsets, g.col_idx)
        t_cc = time.perf_counter() - t
        t = time.perf_counter()
        d16 = oracle.sssp(s16.row_offsets, s16.col_idx, s16.values, 0, quadratic=True)
        t_q = time.perf_counter() - t
    te16 = oracle.teps_edges(s16.row_offsets, d16, INT_MAX)
    log(f"cpu baseline: sssp {t_sssp:.2f}s cc {t_cc:.2f}s, quadratic s16 sssp {t_q:.2f}s (core {pin.core})")
    rec = {"value": te / (t_sssp + t_cc) / 1e9, "unit": "GTEPS", "cores": 1, "kind": "port",
           "sample": f"the full step on the same graph: SSSP from vertex 0 (binary-heap Dijkstra restatement of "
                     f"CPUSSSP, {t_sssp:.2f}s) + CC (CPUCC restatement, {t_cc:.2f}s), one thread pinned to core "
                     f"{pin.core}",
           "pinned_core": pin.core, "cpu": cpu_info(),
           "reference_quadratic_sssp_s16": {"seconds": t_q, "gteps": te16 / t_q / 1e9,
                                            "what": "verbatim O(n^2) CPUSSSP (sssp.cu:127-199) from vertex 0 on "
                                                    "R-MAT scale 16, weights 1..64"}}
    return rec, d_ref, cc_ref, (s16, d16)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--reps", type=int, default=10, help="timed runs of each extra part (median reported)")
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (the oracle check still runs)")
    ap.add_argument("--no-check", action="store_true", help="skip the oracle checks (profiling runs only)")
    ap.add_argument("--parts", default="rmat16,sources,grid,pagerank,batch,serve",
                    help="extra measurements beside the main step: rmat16 (BFS from 0 at scale 16), sources (BFS "
                         "and SSSP from 16 sources at scale 22), grid (BFS/SSSP from the corner of the 2048x2048 "
                         "grid), pagerank (50 iterations on the directed scale-22 R-MAT), batch (256 sources, "
                         "batched vs plain) and serve (Poisson query stream over six resident graphs); '' for none")
    ap.add_argument("--serve-seconds", type=float, default=4.0, help="length of the serving arrival window")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="mean query arrival rate (queries/s)")
    ap.add_argument("--serve-mix", default="0.6,0.3,0.1", help="BFS,SSSP,PPR shares of the query stream")
    ap.add_argument("--serve-graphs", default="", help="comma-separated subset of the six serving graphs")
    ap.add_argument("--serve-config", default="", help="mg_server_config fields, e.g. max_slots=64,thin_lanes=32")
    ap.add_argument("--no-relabel", action="store_true",
                    help="keep the R-MAT graphs in the caller's numbering only (no degree-ordered twin)")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--pmc-pass", action="store_true",
                    help="for counter runs: one warm-up and one measured SSSP, CC and BFS search, then a "
                         "256 MiB device-to-device copy (the calibration kernel); prints per-search stats")
    ap.add_argument("--pmc-pagerank", action="store_true",
                    help="for counter runs: one warm-up and one measured 50-iteration PageRank on the directed "
                         "scale-22 R-MAT (with its twin), then the calibration copy")
    a = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}, WORLD_SIZE {world}); "
              "the engine does not shard a graph over several GPUs", file=sys.stderr)
        return 2

    import mirrorofmapgraph_amd as mg

    mg.check(mg.lib().mg_device_init(0))
    log = (lambda *x: print(*x, file=sys.stderr, flush=True)) if a.verbose else (lambda *x: None)

    if a.pmc_pagerank:
        host = mg.rmat_graph(22, 16, seed=a.seed, symmetric=False)
        with mg.Graph(host, build_csc=True, relabel=not a.no_relabel) as pg:
            o = mg.options(pr_mode=mg.MG_PR_ALL, iter_num=50, pr_tol=1e-6)
            pg.pagerank(o, copy=False)
            _, st = pg.pagerank(o, stats=True, copy=False)
            calibration_copy(256 << 20)
        print(json.dumps({"pagerank": st.as_dict(), "calibration_copy_bytes": 256 << 20}))
        return 0

    t0 = time.perf_counter()
    g = mg.rmat_graph(a.scale, 16, seed=a.seed, weight_max=64)
    log(f"graph s{a.scale}: n={g.n} m={g.m} built in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    # resident with its degree-ordered twin: BFS / SSSP run there (results in the caller's ids)
    dg = mg.Graph(g, symmetric=True, relabel=not a.no_relabel)
    create_s = time.perf_counter() - t0
    log(f"device graph created in {create_s:.1f}s (relabel {not a.no_relabel})")
    src = 0

    if a.pmc_pass:
        out = {}
        for name, fn in [("sssp", lambda: dg.sssp(src, stats=True, copy=False)),
                         ("cc", lambda: dg.cc(stats=True, copy=False)),
                         ("bfs", lambda: dg.bfs(src, stats=True, copy=False))]:
            fn()
            _, st = fn()
            out[name] = st.as_dict()
        nbytes = 256 << 20
        calibration_copy(nbytes)
        out["calibration_copy_bytes"] = nbytes
        print(json.dumps(out))
        dg.close()
        return 0

    # the autotuner's choices for this graph class (mirrorofmapgraph_amd/tuned.json; defaults when absent)
    from mirrorofmapgraph_amd.tuning import tuned_options

    o_sssp, tuned_sssp = tuned_options(g, "sssp")
    o_bfs, tuned_bfs = tuned_options(g, "bfs")
    log(f"tuned options: sssp {tuned_sssp}, bfs {tuned_bfs}")

    def step():
        _, s1 = dg.sssp(src, o_sssp, stats=True, copy=False)
        _, s2 = dg.cc(stats=True, copy=False)
        return s1, s2

    for _ in range(a.warmup):
        s1, s2 = step()
    te_sssp, n_sssp = s1.edges_traversed, s1.vertices_reached
    te_cc = g.m  # every vertex is labelled; its out-edges count once

    sssp_ms, cc_ms = [], []
    mg.check(mg.lib().mg_device_synchronize())
    t = time.perf_counter()
    for _ in range(a.steps):
        s1, s2 = step()
        sssp_ms.append(s1.device_ms)
        cc_ms.append(s2.device_ms)
    mg.check(mg.lib().mg_device_synchronize())
    wall = time.perf_counter() - t
    ms_per_step = 1e3 * wall / a.steps
    value = (te_sssp + te_cc) * a.steps / wall / 1e9

    # BFS from the same source, timed beside the step
    for _ in range(2):
        dg.bfs(src, o_bfs, stats=True, copy=False)
    bfs_ms = []
    for _ in range(a.steps):
        _, sb = dg.bfs(src, o_bfs, stats=True, copy=False)
        bfs_ms.append(sb.device_ms)

    def part(te, ms, nbytes):
        t_ms = float(np.mean(ms))
        gbs = nbytes / (t_ms * 1e-3) / 1e9
        return {"gteps": te / (t_ms * 1e-3) / 1e9, "device_ms": t_ms, "device_ms_median": float(np.median(ms)),
                "device_ms_min": float(np.min(ms)), "algorithmic_bytes": int(nbytes), "achieved_GBs": gbs,
                "hbm_frac": gbs / HBM_PEAK_GBS}

    parts = {
        "sssp": part(te_sssp, sssp_ms, bytes_sssp(te_sssp, n_sssp)) | {"iterations": s1.iterations},
        "cc": part(te_cc, cc_ms, bytes_cc(g.m, g.n)),
        "bfs": part(sb.edges_traversed, bfs_ms, bytes_bfs(sb.edges_traversed, sb.vertices_reached))
               | {"levels": sb.iterations, "bottom_up_levels": sb.bottom_up_levels},
    }
    ss = parts["sssp"]
    traffic = measured_traffic("sssp")
    copy_gbs = copy_bandwidth()
    roofline = {"bound": "hbm", "achieved": ss["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ss["hbm_frac"], "traffic": traffic,
                "kernel": "sssp search (sssp_small + sssp_level_a + sssp_level_b launches of one search)",
                "traffic_source": str(COUNTERS.relative_to(ROOT)) if traffic else None,
                "measured_copy_GBs": copy_gbs, "frac_of_measured_copy": ss["achieved_GBs"] / copy_gbs}

    import oracle

    cpu, d_ref, cc_ref = None, None, None
    try:
        if not a.no_cpu:
            cpu, d_ref, cc_ref, (s16, d16) = cpu_baseline(mg, oracle, g, src, te_sssp + te_cc, log)
            with mg.Graph(s16, symmetric=True) as d16g:
                t16, _ = median_ms(lambda: d16g.sssp(0, stats=True, copy=False), a.reps)
                if not a.no_check:
                    require_equal("rmat16 SSSP vs the quadratic CPUSSSP", d16g.sssp(0), d16)
            cpu["reference_quadratic_sssp_s16"]["gpu_device_ms"] = t16
        if not a.no_check:
            # the step's results, outside the timed loop, against the oracle
            if d_ref is None:
                d_ref = oracle.sssp(g.row_offsets, g.col_idx, g.values, src)
                cc_ref = oracle.cc_union_find(g.row_offsets, g.col_idx)
            require_equal("rmat22 SSSP from 0", dg.sssp(src, o_sssp), d_ref)
            require_equal("rmat22 CC", dg.cc(), cc_ref)
            require_equal("rmat22 BFS from 0", dg.bfs(src, o_bfs), oracle.bfs(g.row_offsets, g.col_idx, src))
            del d_ref, cc_ref
            log("main step checked against the oracle")

        parts_wanted = [p for p in a.parts.split(",") if p]
        if "rmat16" in parts_wanted:
            parts["rmat16_bfs"] = rmat16_part(mg, oracle, a.seed, a.reps, log)
        if "sources" in parts_wanted:
            parts["rmat22_16_sources"] = varied_sources_part(mg, oracle, dg, g, a.seed, log)
        dg.close()
        del g
        if "grid" in parts_wanted:
            parts["grid2048"] = grid_part(mg, oracle, a.seed, a.reps, log)
        if "pagerank" in parts_wanted:
            parts["pagerank_s22d"] = pagerank_part(mg, oracle, a.seed, a.reps, log, relabel=not a.no_relabel)
        if "batch" in parts_wanted:
            parts["batched"] = batched_part(mg, oracle, a.seed, a.reps, log)
        if "serve" in parts_wanted:
            mix = tuple(float(x) for x in a.serve_mix.split(","))
            cfg = {k: int(v) for k, v in (kv.split("=") for kv in a.serve_config.split(",") if kv)}
            parts["serving"] = serving_part(mg, a.seed, a.serve_seconds, a.serve_rate, log, mix,
                                            [x for x in a.serve_graphs.split(",") if x] or None, cfg,
                                            relabel=not a.no_relabel)
    except Mismatch as e:
        print(f"bench.py: PARITY FAILURE: {e}", file=sys.stderr, flush=True)
        return 1

    line = {
        "metric": METRIC, "value": value, "unit": "GTEPS", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic",
        "config": {"workload": "rmat22_sssp_cc", "scale": a.scale, "edge_factor": 16, "rmat": [0.57, 0.19, 0.19, 0.05],
                   "symmetrized": True, "n": int(dg.n), "m": int(dg.m), "weights": [1, 64], "source": src,
                   "seed": a.seed, "parallelism": "single GPU",
                   "layout": "caller ids" if a.no_relabel else "degree-ordered twin (MG_GRAPH_RELABEL)",
                   "graph_create_s": create_s, "tuned_options": {"sssp": tuned_sssp, "bfs": tuned_bfs}},
        "checked_against_oracle": not a.no_check,
        "roofline": roofline, "cpu_baseline": cpu, "parts": parts,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
