"""The reference's application loop, written against include/subway
(tests/cpp/ref_loop.cpp: Graph/GraphPR fields, Subgraph, SubgraphGenerator
with its host post-condition, Partitioner, and the thirteen kernel names of
gpu_kernels.hpp as host launchers), compiled here against the library and
run on graphs partitioned by a small device budget: every application's
output file equals the oracle (PR within 1e-4 of the FP64 oracle)."""
import os
import subprocess

import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


@pytest.fixture(scope="module")
def ref_loop(tmp_path_factory):
    exe = tmp_path_factory.mktemp("refloop") / "ref_loop"
    lib = os.path.join(ROOT, "subway_amd", "lib")
    cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROCM, "include"), os.path.join(ROOT, "tests", "cpp", "ref_loop.cpp"), "-o", str(exe),
           "-L", lib, "-lsubway_amd", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
           f"-Wl,-rpath,{lib}", f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}"]
    subprocess.run(cmd, check=True)
    return str(exe)


APPS = [f"{a}-{m}" for a in ("bfs", "cc", "sssp", "sswp", "pr") for m in ("sync", "async")]


@pytest.mark.parametrize("app", APPS)
def test_reference_loop_on_the_library(sw, gpu, ref_loop, tmp_path, app):
    algo, mode = app.split("-")
    weighted = algo in ("sssp", "sswp")
    g = sw.Graph.rmat(12, 8, seed=21, weighted=weighted, symmetrize=(algo == "cc"))
    esz = 8 if weighted else 4
    budget = 32 << 10
    # every vertex fits the buffer the loop copies its partitions into
    assert g.max_degree <= budget // esz
    path = tmp_path / ("g.bwcsr" if weighted else "g.bcsr")
    g.save(str(path))
    out = tmp_path / "out.txt"
    src = g.max_degree_vertex
    cmd = [ref_loop, app, str(path), str(src), str(budget), str(out)]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = [ln.split() for ln in out.read_text().splitlines()]
    assert len(rows) == g.n
    ptr, adj, w = g.csr()
    if algo == "pr":
        vals = np.array([float(x) for _, x in rows])
        exp, _ = O.pr(ptr, adj, 1e-6)
        assert (np.abs(vals - exp) / exp).max() < 1e-4 + 1e-5      # plus the 6-digit printing
    else:
        vals = np.array([int(x) for _, x in rows], dtype=np.uint64)
        exp = O.expected(algo, ptr, adj, w, src).astype(np.uint64)
        assert np.array_equal(vals, exp), int(np.sum(vals != exp))
    # the budget partitioned the graph (several partitions per iteration)
    assert g.m * esz > 2 * budget
