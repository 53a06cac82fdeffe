"""Renumbered graphs (subway_graph_reorder): the GPU renumbering yields the
CSR a numpy restatement builds (rows in descending in-degree order, ties by
ascending id, each row in its own order with renamed destinations) for every
host layout (unweighted, 8-byte weighted with its packed copy, split,
loaded from a file), and every application, mode, schedule and the
batched runs return the original graph's oracle answers in the original
numbering (bit-exact; PR within 1e-4)."""
import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu
PR_RTOL = 1e-4


def renumbered_csr(ptr, ends, w):
    """The renumbering restated in numpy: perm, iperm and the new CSR."""
    n = len(ptr) - 1
    indeg = np.bincount(ends, minlength=n)
    iperm = np.argsort(-indeg.astype(np.int64), kind="stable").astype(np.int64)
    perm = np.empty(n, np.int64)
    perm[iperm] = np.arange(n)
    deg = np.diff(ptr.astype(np.int64))
    ndeg = deg[iperm]
    nptr = np.zeros(n + 1, np.int64)
    np.cumsum(ndeg, out=nptr[1:])
    idx = np.repeat(ptr[:-1].astype(np.int64)[iperm] - nptr[:-1], ndeg) + np.arange(len(ends), dtype=np.int64)
    ne = perm[ends[idx]].astype(np.uint32)
    nw = None if w is None else w[idx]
    return perm.astype(np.uint32), iperm.astype(np.uint32), nptr.astype(np.uint64), ne, nw


def check_renumbered(g, ptr, ends, w):
    perm, iperm, nptr, ne, nw = renumbered_csr(ptr, ends, w)
    p2, e2, w2 = g.csr()
    gp, gip = g.order()
    assert np.array_equal(gp, perm) and np.array_equal(gip, iperm)
    assert np.array_equal(p2, nptr) and np.array_equal(e2, ne)
    if w is None:
        assert w2 is None
    else:
        assert np.array_equal(w2, nw)
    assert np.array_equal(g.out_degree(), np.diff(nptr.astype(np.int64)).astype(np.uint32))


def expect(algo, ptr, ends, w, src, acc):
    if algo == "pr":
        return O.pr(ptr, ends, acc)[0]
    return O.expected(algo, ptr, ends, w, src)


def same(algo, vals, exp):
    if algo == "pr":
        rel = np.abs(vals.astype(np.float64) - exp) / np.maximum(exp, 1e-30)
        assert rel.max() < PR_RTOL, float(rel.max())
    else:
        bad = np.nonzero(vals != exp)[0]
        assert bad.size == 0, (bad[:10], vals[bad[:10]], exp[bad[:10]])


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("build", ["host", "device"])
def test_renumbering_matches_restatement(sw, gpu, weighted, build):
    g = sw.Graph.rmat(14, 12, seed=3, weighted=weighted, build=build)
    ptr, ends, w = g.csr()
    md, mv = g.max_degree, g.max_degree_vertex
    assert g.order() is None and not g.reordered
    g.reorder()
    assert g.reordered
    check_renumbered(g, ptr, ends, w)
    # the graph's info keeps original ids
    import ctypes
    n, m, wt, d, v = (ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_uint32())
    sw.check(sw.lib.subway_graph_info(g._h, *(ctypes.byref(x) for x in (n, m, wt, d, v))))
    assert (d.value, v.value) == (md, mv)


def test_renumbering_split_layout(sw, gpu):
    g = sw.Graph.rmat(15, 16, seed=8, weighted=True, split_layout=True, build="device")
    ptr, ends, w = g.csr()
    g.reorder()
    check_renumbered(g, ptr, ends, w)
    p, e, w8 = g.split_view()
    p2, e2, w2 = g.csr()
    assert np.array_equal(p, p2) and np.array_equal(e, e2) and np.array_equal(w8.astype(np.uint32), w2)


def test_renumbering_packed_copy_and_file(sw, gpu, tmp_path):
    """A weighted graph whose packed copy exists (a session ran) renumbers
    it too; so does a graph loaded from a file."""
    g = sw.Graph.rmat(14, 16, seed=4, weighted=True, build="host")
    ptr, ends, w = g.csr()
    src = g.max_degree_vertex
    exp = O.sssp(ptr, ends, w, src)
    s = sw.Session(g, "sssp", mode="async", budget=64 << 10)
    vals, st = s.run(src)
    assert st.edge_bytes == 4 and np.array_equal(vals, exp)
    s.close()
    g.reorder()
    check_renumbered(g, ptr, ends, w)
    vals, st = sw.run(g, "sssp", mode="async", source=src, budget=64 << 10)
    assert st.edge_bytes == 4 and np.array_equal(vals, exp)
    path = tmp_path / "g.bwcsr"
    sw.Graph.from_csr(ptr, ends, w).save(str(path))
    f = sw.Graph.load(str(path))
    f.reorder()
    check_renumbered(f, ptr, ends, w)
    vals, st = sw.run(f, "sssp", mode="async", source=src, budget=64 << 10)
    assert np.array_equal(vals, exp)


def test_renumbering_refused_with_a_session(sw, gpu):
    g = sw.Graph.rmat(10, 8, seed=1)
    s = sw.Session(g, "bfs", mode="async")
    with pytest.raises(sw.SubwayError, match="before any session"):
        g.reorder()
    s.close()
    g.reorder()
    with pytest.raises(sw.SubwayError, match="already renumbered"):
        g.reorder()


APPS = [(a, m) for a in ("bfs", "cc", "sssp", "sswp", "pr") for m in ("sync", "async")]


@pytest.mark.parametrize("algo,mode", APPS)
@pytest.mark.parametrize("budget", [0, 96 << 10])
def test_apps_on_renumbered_graph(sw, gpu, algo, mode, budget):
    """Results in the original numbering, equal to the oracle on the
    original graph (CC labels are the smallest original id)."""
    weighted = algo in ("sssp", "sswp")
    g = sw.Graph.rmat(13, 10, seed=11, weighted=weighted, symmetrize=algo == "cc", build="host")
    ptr, ends, w = g.csr()
    src = g.max_degree_vertex if algo != "pr" else 0
    acc = 1e-6
    kw = {"threshold": acc} if algo == "pr" else {}
    g.reorder()
    vals, _ = sw.run(g, algo, mode=mode, source=src, budget=budget, **kw)
    same(algo, vals, expect(algo, ptr, ends, w, src, acc))


@pytest.mark.parametrize("algo", ["sssp", "bfs", "pr", "cc"])
def test_learned_policy_on_renumbered_graph(sw, gpu, algo):
    weighted = algo == "sssp"
    g = sw.Graph.rmat(15, 12, seed=21, weighted=weighted, symmetrize=algo == "cc", build="host")
    ptr, ends, w = g.csr()
    src = g.max_degree_vertex if algo != "pr" else 0
    acc = 1e-6
    kw = {"threshold": acc} if algo == "pr" else {}
    exp = expect(algo, ptr, ends, w, src, acc)
    g.reorder()
    s = sw.Session(g, algo, mode="async", budget=256 << 10, policy="learned", **kw)
    for _ in range(2):   # a second run starts from the same state
        vals, _ = s.run(src)
        same(algo, vals, exp)
    s.close()


def test_resident_and_direction_optimizing_bfs(sw, gpu):
    g = sw.Graph.rmat(14, 16, seed=6, build="host")
    ptr, ends, _ = g.csr()
    exp = O.bfs(ptr, ends, 0)
    g.reorder()
    for push_only in (False, True):
        vals, st = sw.run(g, "bfs", mode="sync", source=0, push_only=push_only)
        assert np.array_equal(vals, exp)
    assert st.pull_rounds == 0


@pytest.mark.parametrize("k", [1, 7, 64])
def test_batched_rows_on_renumbered_graph(sw, gpu, k):
    g = sw.Graph.rmat(13, 12, seed=9, weighted=True, build="host")
    ptr, ends, w = g.csr()
    rng = np.random.default_rng(k)
    srcs = rng.choice(np.nonzero(np.diff(ptr.astype(np.int64)) > 0)[0], k, replace=False).astype(np.uint32)
    g.reorder()
    b = sw.BatchSession(g, "sssp", budget=128 << 10)
    rows, _ = b.run(srcs)
    for i, s in enumerate(srcs):
        assert np.array_equal(rows[i], O.sssp(ptr, ends, w, int(s))), i
    b.close()
