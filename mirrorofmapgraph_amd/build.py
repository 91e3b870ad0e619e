"""Builds libmapgraph_hip.so in-tree for gfx950 (explicit hipcc, no JIT cache).

Usage: python -m mirrorofmapgraph_amd.build [-j N] [--force]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libmapgraph_hip.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
ARCH = os.environ.get("MG_OFFLOAD_ARCH", "gfx950")
DROPIN_SRC = PKG.parent / "tests" / "dropin" / "mapgraph_dropin.cpp"
DROPIN_BIN = PKG.parent / "tests" / "dropin" / "mapgraph_dropin"

# -packed-fp32-ops: no packed FP32 instructions (v_pk_*_f32) are generated.
# -ffp-contract=off: float expressions round as written (PageRank parity).
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off", "-I", str(PKG.parent / "include"),
]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-ffp-contract=off",
              "-I", str(PKG.parent / "include")]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    noisy = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if noisy:
        sys.stderr.write("\n".join(noisy) + "\n")


def sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def build(jobs: int = 8, force: bool = True, verbose: bool = False) -> Path:
    """Compiles every source and links the shared library; returns its path."""
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(exist_ok=True)
    headers = list(CSRC.glob("*.h")) + list((PKG.parent / "include").rglob("*.h"))
    newest_header = max((h.stat().st_mtime for h in headers), default=0.0)
    jobs_list = []
    objs = []
    for src in sources():
        obj = BUILD / (src.name + ".o")
        objs.append(obj)
        if not force and obj.exists() and obj.stat().st_mtime >= max(src.stat().st_mtime, newest_header):
            continue
        if src.suffix == ".hip":
            cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
        else:
            cmd = ["g++", *HOST_FLAGS, "-c", str(src), "-o", str(obj)]
        jobs_list.append(cmd)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for cmd in jobs_list:
            if verbose:
                print(" ".join(cmd))
        list(ex.map(_run, jobs_list))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          "-lpthread"])
    os.replace(tmp, LIB)
    # the command-line driver, linked against the library next to it
    drv = PKG / "driver" / "mapgraph_run.cpp"
    if drv.exists():
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", str(PKG.parent / "include"), str(drv), "-L", str(PKG),
              "-lmapgraph_hip", "-Wl,-rpath,$ORIGIN", "-o", str(PKG / "mapgraph_run")])
    # the C++ drop-in test driver: a MapGraph-style program over the header-only
    # CsrProblem / EnactorVertexCentric API (compiled by hipcc: the enactor's kernels)
    dropin = DROPIN_SRC
    if dropin.exists():
        _run([HIPCC, *[f for f in HIP_FLAGS if f != "-fPIC"], str(dropin), "-o", str(DROPIN_BIN)])
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(jobs=a.j, force=not a.incremental, verbose=a.v))


if __name__ == "__main__":
    main()
