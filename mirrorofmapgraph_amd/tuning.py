"""Tuned engine parameters per graph class, and the command that produces them.

mg_autotune (include/mapgraph_c.h) picks, for one graph and one search kind, the
mg_options fields that decide how a search maps onto the GPU.  Graphs of one class
(same size and shape) tune alike, so results are kept in tuned.json next to this
file, keyed by graph_class() and the search kind, and applied with
tuned_options().  The table is produced on the GPU by

    python -m mirrorofmapgraph_amd.tuning --out <path>

which tunes the benchmark graphs (R-MAT scales 16, 20, 22 and the 2048 x 2048
grid) for their searches and writes the table; tuned.json holds the output of such
a run.  Tuned options never change results (mg_autotune checks every candidate
bit for bit), only the time.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
from pathlib import Path

import numpy as np

from . import TUNABLE, HostCSR, options

TABLE = Path(__file__).resolve().parent / "tuned.json"


def graph_class(host: HostCSR, symmetric: bool = True) -> str:
    """Class key: log2 of n and m, hub-shaped or bounded-degree, weight range, symmetry."""
    deg = host.degrees()
    n, m = host.n, host.m
    hub = n > 0 and int(deg.max()) >= 1024 and deg.max() >= 32 * (m / max(n, 1))
    w = f"w{int(host.values.min())}-{int(host.values.max())}" if host.values is not None and m else "w1"
    return (f"v{round(math.log2(max(n, 1)))}-e{round(math.log2(max(m, 1)))}-{'hub' if hub else 'flat'}-{w}-"
            f"{'sym' if symmetric else 'dir'}")


def load_table(path: Path | None = None) -> dict:
    p = path or TABLE
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def tuned_options(host: HostCSR, what: str, base=None, symmetric: bool = True, table: dict | None = None):
    """(options, applied): base (or the defaults) with the table's tuned fields for this graph class and
    search kind; applied is the dict of fields taken from the table ({} when the class is not in it)."""
    o = base if base is not None else options()
    entry = (table if table is not None else load_table()).get(graph_class(host, symmetric), {}).get(what, {})
    applied = {k: v for k, v in entry.get("options", {}).items() if k in TUNABLE}
    for k, v in applied.items():
        setattr(o, k, int(v))
    return o, applied


def _tune_all(log) -> dict:
    import mirrorofmapgraph_amd as mg

    mg.check(mg.lib().mg_device_init(0))
    table = {}
    # R-MAT graphs are tuned as the benchmark keeps them: with the degree-ordered twin
    plans = [
        ("rmat16", lambda: mg.rmat_graph(16, 16, seed=42, weight_max=64), ["bfs", "sssp"], 8, 64, True),
        ("rmat20", lambda: mg.rmat_graph(20, 16, seed=42, weight_max=64), ["bfs", "sssp", "bfs_batch", "sssp_batch"], 8,
         64, False),
        ("rmat22", lambda: mg.rmat_graph(22, 16, seed=42, weight_max=64), ["bfs", "sssp"], 8, 64, True),
        ("grid2048", lambda: mg.grid_graph(2048, seed=42, weight_max=16), ["bfs", "sssp", "bfs_batch", "sssp_batch"], 3,
         32, False),
    ]
    for name, make, kinds, nsingle, nbatch, relabel in plans:
        host = make()
        key = graph_class(host)
        srcs = mg.pick_sources(host, max(nsingle, nbatch), seed=7)
        single = np.concatenate([[0], srcs[: nsingle - 1]]).astype(np.int32)
        with mg.Graph(host, symmetric=True, relabel=relabel) as dg:
            for what in kinds:
                s = srcs[:nbatch] if what.endswith("_batch") else single
                best, rep = dg.autotune(what, s, reps=3)
                chosen = {k: getattr(best, k) for k in TUNABLE}
                table.setdefault(key, {})[what] = {"options": chosen, "report": rep, "graph": name,
                                                   "sources": len(s), "relabel": relabel}
                log(f"{name} {what}: {rep['start_ms']:.3f} -> {rep['best_ms']:.3f} ms over {rep['candidates']} "
                    f"settings: {chosen}")
    return table


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(TABLE), help="where to write the table")
    a = ap.parse_args()
    table = _tune_all(lambda *x: print(*x, flush=True))
    Path(a.out).write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")
    print(f"wrote {a.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
