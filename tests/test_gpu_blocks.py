"""Block-compressed rows (subway_graph_build_blocks) on the GPU.

The layout: after build_blocks every host layout of the graph holds its rows
in ascending destination order (a stable sort: tests/_blocks.sort_rows) and
the block words and row offsets equal the Python restatement's
(tests/_blocks.encode) word for word, for unweighted graphs, weighted 8-byte
lists (with their packed copy), weights wider than a byte, split-layout and
renumbered graphs.

The runs: every application and mode on graphs that move whole id ranges in
the block layout (partitioned budgets, one and two buffers, the resident
graph, the learned schedule with its whole ranges and in-place visits)
returns the oracle's answer on the original graph (bit-exact; PR within
1e-4), and the same values as sessions that move the plain edges."""
import zlib

import numpy as np
import pytest

from tests import _blocks as B
from tests import _oracle as O

pytestmark = pytest.mark.gpu
PR_RTOL = 1e-4
APPS = [(a, m) for a in ("bfs", "cc", "sssp", "sswp", "pr") for m in ("sync", "async")]


def check_layout(g, ptr, ends, w):
    se, sw8 = B.sort_rows(ptr, ends, w)
    g.build_blocks()
    p2, e2, w2 = g.csr()
    assert np.array_equal(p2, ptr) and np.array_equal(e2, se)
    if w is not None:
        assert np.array_equal(w2, sw8)
    row_ptr, words, wbits, wbase = g.blocks_view()
    er, ew, eb, ebase = B.encode(ptr, se, sw8)
    assert (wbits, wbase) == (eb, ebase)
    assert np.array_equal(row_ptr, er)
    assert np.array_equal(words, ew)
    return g


def edges_graph(sw, rng, n, m, wmax=None, hub=None):
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    if hub is not None:
        v, d = hub
        src = np.concatenate([src, np.full(d, v, np.uint32)])
        dst = np.concatenate([dst, rng.integers(0, n, d).astype(np.uint32)])
    w = None if wmax is None else rng.integers(1, wmax + 1, len(src)).astype(np.uint32)
    return sw.Graph.from_edges(n, src, dst, w)


@pytest.mark.parametrize("kind", ["unweighted", "weighted", "wide-weights", "split", "renumbered", "empty-rows"])
def test_layout_matches_restatement(sw, gpu, kind):
    rng = np.random.default_rng(zlib.crc32(kind.encode()))
    if kind == "unweighted":
        g = edges_graph(sw, rng, 3000, 20000, hub=(5, 3000))
    elif kind == "weighted":
        g = edges_graph(sw, rng, 3000, 20000, wmax=64, hub=(7, 2000))
    elif kind == "wide-weights":
        g = edges_graph(sw, rng, 2000, 15000, wmax=100000)
    elif kind == "split":
        g = sw.Graph.rmat(11, 16, seed=3, weighted=True, split_layout=True)
    elif kind == "renumbered":
        g = sw.Graph.rmat(12, 8, seed=4, weighted=True)
        g.reorder()
    else:
        g = edges_graph(sw, rng, 5000, 300, wmax=64)       # mostly rows of degree 0
    ptr, ends, w = g.csr()
    check_layout(g, ptr, ends, w)


def test_build_twice_and_after_session_is_refused(sw, gpu):
    g = sw.Graph.rmat(8, 8, seed=1, weighted=True)
    g.build_blocks()
    with pytest.raises(sw.SubwayError):
        g.build_blocks()
    h = sw.Graph.rmat(8, 8, seed=1, weighted=True)
    s = sw.Session(h, "sssp", mode="async")
    with pytest.raises(sw.SubwayError):
        h.build_blocks()
    s.close()


def expected(algo, ptr, adj, w, src):
    if algo == "pr":
        return O.pr(ptr, adj, 1e-6)[0]
    return O.expected(algo, ptr, adj, w, src)


def check(algo, vals, exp):
    if algo == "pr":
        rel = np.abs(vals.astype(np.float64) - exp) / np.maximum(exp, 1e-30)
        assert rel.max() < PR_RTOL, float(rel.max())
    else:
        bad = np.flatnonzero(vals != exp)
        assert bad.size == 0, (bad[:10], vals[bad[:10]], exp[bad[:10]])


@pytest.mark.parametrize("algo,mode", APPS)
@pytest.mark.parametrize("budget,buffers", [(96 << 10, 2), (96 << 10, 1), (8 << 20, 2)])
def test_apps_on_blocks(sw, gpu, algo, mode, budget, buffers):
    """Partitioned budgets move fully active contiguous partitions (every
    vertex is active in an async run's first iteration) in the block
    layout; 8 MiB holds the whole graph (resident, blocked)."""
    weighted = algo in ("sssp", "sswp")
    g = sw.Graph.rmat(14, 16, seed=zlib.crc32(algo.encode()) & 0xFFFF, weighted=weighted,
                      symmetrize=(algo == "cc"))
    ptr, adj, w = g.csr()                    # the original rows: results must not change
    src = g.max_degree_vertex if algo != "cc" else 0
    exp = expected(algo, ptr, adj, w, src)
    g.build_blocks()
    kw = {"threshold": 1e-6} if algo == "pr" else {}
    s = sw.Session(g, algo, mode=mode, budget=budget, buffers=buffers, **kw)
    for _ in range(2):
        v, st = s.run(src)
        check(algo, v, exp)
        if mode == "async" or budget == 8 << 20:
            assert st.blocked_units > 0 and st.h2d_blocked_bytes > 0
            assert st.h2d_blocked_bytes < st.blocked_edges * st.edge_bytes
    s.close()
    vp, _ = sw.Session(g, algo, mode=mode, budget=budget, buffers=buffers, blocks=False, **kw).run(src)
    if algo != "pr":
        assert np.array_equal(v, vp)


@pytest.mark.parametrize("algo", ["bfs", "sssp", "sswp", "cc", "pr"])
def test_learned_schedule_on_blocks(sw, gpu, algo):
    """The learned schedule's whole ranges and in-place visits of resident
    ranges read blocked slices."""
    from subway_amd import policy
    weighted = algo in ("sssp", "sswp")
    g = sw.Graph.rmat(15, 16, seed=9, weighted=weighted, symmetrize=(algo == "cc"))
    ptr, adj, w = g.csr()
    src = g.max_degree_vertex if algo != "cc" else 0
    exp = expected(algo, ptr, adj, w, src)
    g.build_blocks()
    kw = {"threshold": 1e-6} if algo == "pr" else {}
    for params in (policy.default_params(), None):
        s = sw.Session(g, algo, mode="async", budget=int(g.m * 4 * 0.6), policy="learned", params=params,
                       theta=0.05, resident_min_bytes=1, **kw)
        v, st = s.run(src)
        s.close()
        check(algo, v, exp)
        assert st.blocked_units > 0


@pytest.mark.parametrize("algo", ["bfs", "sssp"])
def test_hub_and_gathered_units_beside_blocks(sw, gpu, algo):
    """A vertex larger than a buffer streams in plain chunks and compacted
    units are gathered as plain edges, beside blocked partitions, in one
    run."""
    rng = np.random.default_rng(21)
    g = edges_graph(sw, rng, 20000, 160000, wmax=64 if algo == "sssp" else None, hub=(17, 30000))
    ptr, adj, w = g.csr()
    exp = O.expected(algo, ptr, adj, w, 17)
    g.build_blocks()
    for transfer in ("auto", "host", "zerocopy"):
        v, st = sw.run(g, algo, mode="async", source=17, budget=64 << 10, transfer=transfer)
        check(algo, v, exp)
        if transfer != "zerocopy":
            assert st.blocked_units > 0


def test_direction_optimizing_bfs_on_blocks(sw, gpu):
    g = sw.Graph.rmat(16, 16, seed=0)
    ptr, adj, _ = g.csr()
    exp = O.bfs(ptr, adj, 0)
    g.build_blocks()
    for mode in ("sync", "async"):
        v, st = sw.run(g, "bfs", mode=mode, source=0)
        assert np.array_equal(v, exp)
        assert st.blocked_units == 1 and st.pull_rounds > 0


def test_blocks_on_renumbered_scale18_match_plain(sw, gpu):
    """A mid-size renumbered weighted R-MAT: blocked runs equal plain runs
    and Dijkstra, and move fewer bytes."""
    g = sw.Graph.rmat(18, 16, seed=42, weighted=True)
    src = g.max_degree_vertex                # an original id: runs take and return original ids
    g.reorder()
    plain = sw.Session(g, "sssp", mode="async", budget=8 << 20)
    vp, sp = plain.run(src)
    plain.close()
    g2 = sw.Graph.rmat(18, 16, seed=42, weighted=True)
    g2.reorder()
    g2.build_blocks()
    s = sw.Session(g2, "sssp", mode="async", budget=8 << 20)
    vb, sb = s.run(src)
    s.close()
    assert np.array_equal(vp, vb)
    ptr, adj, w = sw.Graph.rmat(18, 16, seed=42, weighted=True).csr()
    assert np.array_equal(vb, O.sssp(ptr, adj, w, src))
    assert sb.blocked_units > 0 and sb.h2d_bytes < 0.8 * sp.h2d_bytes


@pytest.mark.parametrize("algo,mode", [("sssp", "async"), ("sswp", "async"), ("sssp", "sync"), ("bfs", "async")])
def test_unpinned_split_arrays_gather_on_the_host(sw, gpu, algo, mode):
    """A split-layout graph whose split arrays moved to ordinary memory (as a
    graph at the pinned-memory limit does by itself): whole ranges move as
    blocks, compacted units and hub chunks through host gathers / pageable
    DMA, with the oracle's answers."""
    rng = np.random.default_rng(zlib.crc32(f"{algo}{mode}".encode()))
    n, m = 20000, 160000
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    src = np.concatenate([src, np.full(30000, 17, np.uint32)])
    dst = np.concatenate([dst, rng.integers(0, n, 30000).astype(np.uint32)])
    w = rng.integers(1, 65, len(src)).astype(np.uint32)
    g0 = sw.Graph.from_edges(n, src, dst, w)
    ptr, adj, ww = g0.csr()
    exp = O.expected(algo, ptr, adj, ww, 17)
    path = None
    import tempfile, os
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "g.bwcsr")
        g0.save(path)
        g = sw.Graph.load(path, split_layout=True)
    g.split_view()                          # the split arrays exist
    g.build_blocks(unpin_split=True)
    for transfer in ("auto", "host"):
        v, st = sw.run(g, algo, mode=mode, source=17, budget=64 << 10, transfer=transfer)
        check(algo, v, exp)
        assert st.edge_bytes == 5
        if mode == "async":
            assert st.blocked_units > 0


def test_row_offsets_end_at_the_word_count(sw, gpu):
    """Regression: the word count is read back only after the kernel that
    computes the row offsets has finished (a read that did not wait for the
    build's stream sized the words short).  Scale 22: offsets are monotone,
    the last equals the word count, and decoding sample rows gives the
    sorted CSR."""
    g = sw.Graph.rmat(22, 16, seed=42, weighted=True)
    g.reorder()
    g.build_blocks()
    rp, words, wbits, wbase = g.blocks_view()
    assert int(rp[-1]) == len(words)
    assert np.all(np.diff(rp.astype(np.int64)) >= 0)
    ptr, ends, w = g.csr()
    deg = np.diff(ptr.astype(np.int64))
    rows = np.concatenate([np.argsort(-deg)[:3], np.arange(g.n - 5, g.n), [g.n // 2, g.n // 3]])
    for v in rows:
        v = int(v)
        a, b = int(ptr[v]), int(ptr[v + 1])
        r0, r1 = int(rp[v]), int(rp[v + 1])
        sub_ptr = np.array([0, b - a], np.uint64)
        sub_rp = np.array([0, r1 - r0], np.uint64)
        de, dw = B.decode(sub_rp, words[r0:r1], sub_ptr, wbits, wbase, True)
        assert np.array_equal(de, ends[a:b]) and np.array_equal(dw, w[a:b]), v


@pytest.mark.parametrize("algo,mode", APPS)
def test_gathered_units_move_block_rows(sw, gpu, algo, mode):
    """Units gathered from the active list: the GPU reads the block rows of
    their active vertices from the mapped block words (transfer="zerocopy":
    no DMA at all, so every blocked byte was gathered), with the oracle's
    answers, and the same values as plain gathers of the edges."""
    weighted = algo in ("sssp", "sswp")
    g = sw.Graph.rmat(14, 16, seed=zlib.crc32(f"g{algo}".encode()) & 0xFFFF, weighted=weighted,
                      symmetrize=(algo == "cc"))
    ptr, adj, w = g.csr()
    src = g.max_degree_vertex if algo != "cc" else 0
    exp = expected(algo, ptr, adj, w, src)
    g.build_blocks()
    kw = {"threshold": 1e-6} if algo == "pr" else {}
    for transfer in ("zerocopy", "auto"):
        s = sw.Session(g, algo, mode=mode, budget=96 << 10, transfer=transfer, **kw)
        for _ in range(2):
            v, st = s.run(src)
            check(algo, v, exp)
            if transfer == "zerocopy":
                assert st.h2d_dma_bytes == 0
                assert st.blocked_units > 0 and 0 < st.h2d_blocked_bytes <= st.h2d_gather_bytes
        s.close()
    vp, _ = sw.Session(g, algo, mode=mode, budget=96 << 10, transfer="zerocopy", blocks=False, **kw).run(src)
    if algo != "pr":
        assert np.array_equal(v, vp)
