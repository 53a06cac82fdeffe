"""The five benchmark configurations at their stated sizes, on the GPU.

Collected after the small-graph parity tests: each builds its R-MAT input on
the device (the host generator's CSR, byte for byte; test_gpu_rmat.py), runs
the configuration, checks it, and frees the graph before the next one.

  c1  BFS-sync, R-MAT scale 16, no budget cap      -> test_gpu_parity.py
  c2  SSSP-async + SSWP-async, weighted scale 24, 1 GiB budget: both
      policies bit-exact against Dijkstra / widest-path Dijkstra
  c3  CC-async on the symmetrised scale 25 (1.07 G edges), 2 GiB: bit-exact
      against the min-ancestor labelling; PR-async on scale 25 at tau = 1e-6,
      2 GiB: plain and learned schedules within 1e-4 of each other and both
      satisfying the PageRank fixpoint to the threshold (the FP64 oracle
      takes minutes at this size; it pins PR up to scale 20 elsewhere)
  c4  64-source batched SSSP, weighted scale 26, 4 GiB, autotuned: every row
      equal to a single-source run of its source, row 0 to Dijkstra, three
      rows to the OpenMP build
  c5  SSSP-async and BFS-async, weighted scale 28 edge factor 32 (8.59 G
      edges, split 5-byte transfers), 16 GiB: learned bit-exact with plain
      and with the OpenMP build of the same rules; then the graph is
      renumbered by in-degree and its rows block-compressed (the headline
      layout), and both policies give the same answers again
and a 17 GB .bwcsr64 file loaded straight into its pinned layout.

Oracles and OpenMP runs go to worker threads (ctypes drops the GIL) while
the GPU runs, so the CPU checks overlap the device work.
"""
import gc
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu
PR_RTOL = 1e-4
GiB = 1 << 30


def log(msg):
    print(f"[full-size] {msg}", flush=True)


def _shipped_params():
    from subway_amd import policy
    return policy.default_params()


def _free(*objs):
    for o in objs:
        close = getattr(o, "close", None) or getattr(o, "free", None)
        if close:
            close()
    gc.collect()


def _run(sw, g, algo, policy, src, budget, **kw):
    params = _shipped_params() if policy == "learned" else None
    s = sw.Session(g, algo, mode="async", budget=budget, policy=policy, params=params, **kw)
    t = time.perf_counter()
    vals, st = s.run(src)
    log(f"{algo}-async {policy}: {1e3 * (time.perf_counter() - t):.1f} ms, {st.global_iterations} iterations, "
        f"{st.partitions} partitions, H2D {st.h2d_bytes / 1e9:.2f} GB")
    s.close()
    return vals, st


@pytest.mark.timeout(600)
def test_c2_sssp_sswp_async_scale24(sw, gpu):
    t0 = time.perf_counter()
    g = sw.Graph.rmat(24, 16, seed=42, weighted=True)
    assert g.n == 1 << 24 and g.m > 268_000_000
    log(f"c2 graph: {g.n} vertices, {g.m} edges in {time.perf_counter() - t0:.1f} s")
    ptr, adj, w = g.csr()
    src = g.max_degree_vertex
    with ThreadPoolExecutor(3) as ex:
        exp = {"sssp": ex.submit(O.sssp, ptr, adj, w, src), "sswp": ex.submit(O.sswp, ptr, adj, w, src)}
        omp = ex.submit(O.omp_label, "sssp", ptr, adj, w, src)
        got = {}
        for algo in ("sssp", "sswp"):
            for policy in ("plain", "learned"):
                vals, st = _run(sw, g, algo, policy, src, GiB)
                assert st.partitions > st.global_iterations      # the 1 GiB budget partitions the graph
                got[(algo, policy)] = vals
        for algo in ("sssp", "sswp"):
            e = exp[algo].result()
            for policy in ("plain", "learned"):
                bad = np.flatnonzero(got[(algo, policy)] != e)
                assert bad.size == 0, (algo, policy, bad[:10])
        assert np.array_equal(omp.result()[0], exp["sssp"].result())
    log(f"c2 done in {time.perf_counter() - t0:.1f} s")
    _free(g)


@pytest.mark.timeout(600)
def test_c3_cc_async_symmetric_scale25(sw, gpu):
    t0 = time.perf_counter()
    g = sw.Graph.rmat(25, 16, seed=42, symmetrize=True)
    assert g.m > 1_000_000_000
    log(f"c3 CC graph: {g.n} vertices, {g.m} edges in {time.perf_counter() - t0:.1f} s")
    ptr, adj, _ = g.csr()
    with ThreadPoolExecutor(1) as ex:
        exp = ex.submit(O.cc, ptr, adj)
        plain, st = _run(sw, g, "cc", "plain", 0, 2 * GiB)
        assert st.partitions > st.global_iterations
        learned, _ = _run(sw, g, "cc", "learned", 0, 2 * GiB)
        e = exp.result()
    assert np.array_equal(plain, e) and np.array_equal(learned, e)
    log(f"c3 CC done in {time.perf_counter() - t0:.1f} s")
    _free(g)


def pr_fixpoint_residual(ptr, adj, x):
    """r[v] = 0.15 + 0.85 * sum_{u->v} x[u]/deg(u) - x[v], and the magnitude
    the FP32 results are rounded at.  Delta-push keeps x + delta equal to
    0.15 + 0.85 P^T x, so at its end r is the final residual: in [0, tau] up
    to the rounding of the FP32 shares and results."""
    n = len(x)
    deg = np.diff(ptr).astype(np.int64)
    share = np.zeros(n, np.float64)
    nz = deg > 0
    share[nz] = x[nz].astype(np.float64) / deg[nz]
    src_of = np.repeat(np.arange(n, dtype=np.uint32), deg)
    y = np.bincount(adj, weights=share[src_of], minlength=n)
    del src_of
    x64 = x.astype(np.float64)
    return 0.15 + 0.85 * y - x64, x64 + 0.85 * y


@pytest.mark.timeout(600)
def test_c3_pr_async_scale25(sw, gpu):
    t0 = time.perf_counter()
    g = sw.Graph.rmat(25, 16, seed=42)
    log(f"c3 PR graph: {g.n} vertices, {g.m} edges in {time.perf_counter() - t0:.1f} s")
    tau = 1e-6
    plain, _ = _run(sw, g, "pr", "plain", 0, 2 * GiB, threshold=tau)
    learned, _ = _run(sw, g, "pr", "learned", 0, 2 * GiB, threshold=tau)
    rel = np.abs(plain.astype(np.float64) - learned) / np.maximum(plain.astype(np.float64), 1e-30)
    assert rel.max() < PR_RTOL, float(rel.max())
    ptr, adj, _ = g.csr()
    for name, x in (("plain", plain), ("learned", learned)):
        r, mag = pr_fixpoint_residual(ptr, adj, x)
        tol = 1e-6 * mag + 1e-9
        lo, hi = int(np.argmin(r + tol)), int(np.argmax(r - tol))
        assert r[lo] >= -tol[lo], (name, lo, float(r[lo]), float(mag[lo]))
        assert r[hi] <= tau + tol[hi], (name, hi, float(r[hi]), float(mag[hi]))
        log(f"PR {name}: fixpoint residual in [{r.min():.3g}, {r.max():.3g}], mass {float(x.sum()):.6g}")
    log(f"c3 PR done in {time.perf_counter() - t0:.1f} s")
    _free(g)


@pytest.mark.timeout(600)
def test_c4_batched_64_sources_scale26(sw, gpu):
    t0 = time.perf_counter()
    g = sw.Graph.rmat(26, 16, seed=42, weighted=True)
    log(f"c4 graph: {g.n} vertices, {g.m} edges in {time.perf_counter() - t0:.1f} s")
    ptr, adj, w = g.csr()
    deg = np.diff(ptr)
    rng = np.random.default_rng(42)
    sources = np.sort(rng.choice(np.flatnonzero(deg >= 1), size=64, replace=False)).astype(np.uint32)
    with ThreadPoolExecutor(2) as ex:
        oracle_row = ex.submit(O.sssp, ptr, adj, w, int(sources[0]))
        omp_rows = ex.submit(lambda: [O.omp_label("sssp", ptr, adj, w, int(s))[0] for s in sources[1:4]])
        bs = sw.BatchSession(g, "sssp", budget=4 * GiB)
        best, times = bs.tune(sources)
        log(f"autotuner: unroll {best}, runs {[round(t) for t in times]} ms")
        rows, st = bs.run(sources)
        bs.close()
        assert st.partitions > st.global_iterations
        seq = sw.Session(g, "sssp", mode="async", budget=4 * GiB)
        for i, s in enumerate(sources):
            v, _ = seq.run(int(s), copy=False)
            assert np.array_equal(v, rows[i]), (i, int(s), int(np.sum(v != rows[i])))
        seq.close()
        log(f"64 rows equal to 64 single-source runs ({time.perf_counter() - t0:.1f} s)")
        for i, r in enumerate(omp_rows.result()):
            assert np.array_equal(rows[i + 1], r), i + 1
        assert np.array_equal(rows[0], oracle_row.result())
    log(f"c4 done in {time.perf_counter() - t0:.1f} s")
    del rows
    _free(g)


@pytest.mark.timeout(600)
def test_large_file_loads_into_pinned_layouts(sw, gpu, tmp_path):
    """An 18 GB .bwcsr64 file (weighted R-MAT scale 26, edge factor 33: 2.2 G
    edges, an 8-byte list of 17.7 GB, above the 16 GiB that picks the split
    layout; edge factor 32 drops just under it once self-loops go) is read by parallel positioned
    reads straight into the pinned split arrays, never into an 8-byte list,
    and since its ids and weight span fit one word a packed copy is built from
    them: the loaded graph equals the generated one and runs with 4-byte
    transfers to the same distances."""
    import shutil
    g = sw.Graph.rmat(26, 33, seed=42, weighted=True)
    assert g.m * 8 >= 16 * GiB
    need = 16 + 8 * (g.n + g.m)
    if shutil.disk_usage(str(tmp_path)).free < 2 * need:
        pytest.skip(f"needs {2 * need / 1e9:.0f} GB of free disk for the file")
    path = tmp_path / "g.bwcsr64"
    t0 = time.perf_counter()
    g.save(str(path))
    t1 = time.perf_counter()
    h = sw.Graph.load(str(path))
    t2 = time.perf_counter()
    log(f"scale 26 ef 33: {g.m} edges, save {t1 - t0:.1f} s, load {t2 - t1:.1f} s "
        f"({path.stat().st_size / (t2 - t1) / 1e9:.1f} GB/s)")
    assert (h.n, h.m, h.max_degree, h.max_degree_vertex) == (g.n, g.m, g.max_degree, g.max_degree_vertex)
    assert np.array_equal(h.node_pointer(), g.node_pointer())
    hp, he, hw = h.split_view()
    _, ge, gw = g.csr()
    assert np.array_equal(he, ge) and np.array_equal(hw.astype(np.uint32), gw)
    del ge, gw
    src = g.max_degree_vertex
    a, sa = _run(sw, g, "sssp", "plain", src, 4 * GiB)
    b, sb = _run(sw, h, "sssp", "plain", src, 4 * GiB)
    assert np.array_equal(a, b) and sa.edge_bytes == sb.edge_bytes == 4
    del hp, he, hw
    path.unlink()
    _free(g, h)


@pytest.mark.timeout(600)
def test_c5_learned_vs_plain_scale28(sw, gpu):
    t0 = time.perf_counter()
    g = sw.Graph.rmat(28, 32, seed=42, weighted=True)
    assert g.m > 8_589_000_000
    log(f"c5 graph: {g.n} vertices, {g.m} edges in {time.perf_counter() - t0:.1f} s")
    ptr, adj, w8 = g.split_view()          # the pinned split arrays, no copy
    src = g.max_degree_vertex
    with ThreadPoolExecutor(1) as ex:
        omp = {a: ex.submit(lambda a=a: O.omp_label(a, ptr, adj, w8 if a == "sssp" else None, src)[0])
               for a in ("sssp", "bfs")}
        for algo in ("sssp", "bfs"):
            plain, st = _run(sw, g, algo, "plain", src, 16 * GiB)
            assert st.partitions > st.global_iterations and st.edge_bytes == 5
            learned, _ = _run(sw, g, algo, "learned", src, 16 * GiB)
            bad = np.flatnonzero(plain != learned)
            assert bad.size == 0, (algo, bad[:10])
            e = omp[algo].result()
            bad = np.flatnonzero(plain != e)
            assert bad.size == 0, (algo, bad[:10], plain[bad[:10]], e[bad[:10]])
            del plain, learned
        exp = {a: omp[a].result() for a in ("sssp", "bfs")}
    del ptr, adj, w8                         # views of the split arrays the block build rewrites
    # the headline layout: renumbered by in-degree, rows block-compressed
    t1 = time.perf_counter()
    g.reorder()
    g.build_blocks(unpin_split=True)
    log(f"c5 renumbered and block layout built in {time.perf_counter() - t1:.1f} s")
    for algo in ("sssp", "bfs"):
        plain, st = _run(sw, g, algo, "plain", src, 16 * GiB)
        assert st.blocked_units > 0 and st.h2d_bytes < 0.7 * st.edge_bytes * g.m
        learned, _ = _run(sw, g, algo, "learned", src, 16 * GiB)
        for name, v in (("plain", plain), ("learned", learned)):
            bad = np.flatnonzero(v != exp[algo])
            assert bad.size == 0, (algo, name, bad[:10])
        del plain, learned
    log(f"c5 done in {time.perf_counter() - t0:.1f} s")
    _free(g)
