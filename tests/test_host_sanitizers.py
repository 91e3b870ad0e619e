"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

tests/sanitize/host_tools_check.cpp drives the host half of the library
(mirrorofmapgraph_amd/csrc/graph_host.cpp: Matrix Market reader with malformed
inputs, R-MAT and grid generators, CSR / CSC builders under every flag
combination, hashed weights, the source picker and the argument checks) and the
CPU oracle (oracle/mg_oracle.c) on fixtures, generated, empty and tiny graphs.
Both are compiled here with -fsanitize=address,undefined (host compilers only, no
device code) and any sanitizer report aborts the run."""
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

SAN = ["-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="needs gcc/g++")
def test_host_tools_and_oracle_under_asan_ubsan(tmp_path):
    inc = str(ROOT / "include")
    steps = [
        ["gcc", *SAN, "-fopenmp", "-ffp-contract=off", "-c", str(ROOT / "oracle" / "mg_oracle.c"), "-o",
         str(tmp_path / "oracle.o")],
        ["g++", *SAN, "-std=c++17", "-I", inc, "-c", str(ROOT / "mirrorofmapgraph_amd" / "csrc" / "graph_host.cpp"),
         "-o", str(tmp_path / "graph_host.o")],
        ["g++", *SAN, "-std=c++17", "-I", inc, str(ROOT / "tests" / "sanitize" / "host_tools_check.cpp"),
         str(tmp_path / "graph_host.o"), str(tmp_path / "oracle.o"), "-fopenmp", "-lpthread", "-o",
         str(tmp_path / "check")],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, " ".join(cmd) + "\n" + r.stderr[-4000:]
    scratch = tmp_path / "scratch"
    scratch.mkdir()
    env = {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=1", "UBSAN_OPTIONS": "print_stacktrace=1",
           "OMP_NUM_THREADS": "4", "PATH": "/usr/bin:/bin"}
    r = subprocess.run([str(tmp_path / "check"), str(GOLDEN), str(scratch)], capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-8000:]
    assert "0 failure(s)" in r.stdout
