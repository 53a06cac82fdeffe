"""Early read-back of single-source results (subway_options.early_readback).

A run may copy its values to the host while its last iterations still run,
then copy only the values that changed after that copy and patch them in.
Whatever happened (RunStats.early_readback: 0 copied at the end, 1 copied
early and patched, 2 copied early and then again in full), the host values
are the oracle's, for every label application, in both modes, on plain and
renumbered graphs, over repeated runs of one session."""
import zlib

import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu
LABEL_APPS = [(a, m) for a in ("bfs", "cc", "sssp", "sswp") for m in ("sync", "async")]


def graph(sw, algo, scale, ef, renumber):
    g = sw.Graph.rmat(scale, ef, seed=zlib.crc32(f"{algo}{scale}".encode()) & 0xFFFF,
                      weighted=algo in ("sssp", "sswp"), symmetrize=algo == "cc")
    ptr, adj, w = g.csr()
    src = 0 if algo == "cc" else g.max_degree_vertex
    exp = O.expected(algo, ptr, adj, w, src)
    if renumber:
        g.reorder()
    return g, src, exp


@pytest.mark.parametrize("renumber", [False, True])
@pytest.mark.parametrize("algo,mode", LABEL_APPS)
def test_early_copy_then_patch(sw, gpu, algo, mode, renumber):
    """Forced early copy (after the first iteration) with a patch list large
    enough for every vertex: the patched host values equal the oracle's, and
    equal those of a session that copies at the end; two runs in a row."""
    g, src, exp = graph(sw, algo, 14, 8, renumber)
    s = sw.Session(g, algo, mode=mode, budget=96 << 10, early_readback=g.n)
    for _ in range(2):
        v, st = s.run(src)
        assert np.array_equal(v, exp)
        if st.global_iterations > 1:
            assert st.early_readback == 1
        assert st.values_patched <= g.n
    s.close()
    v0, st0 = sw.Session(g, algo, mode=mode, budget=96 << 10, early_readback=-1).run(src)
    assert st0.early_readback == 0 and np.array_equal(v0, exp)


@pytest.mark.parametrize("algo,mode", [("bfs", "sync"), ("sssp", "sync"), ("sssp", "async")])
def test_patch_list_overflow_copies_in_full(sw, gpu, algo, mode):
    """A patch list of one value: a run whose values change after the early
    copy copies them again in full, with the same answer."""
    g, src, exp = graph(sw, algo, 13, 8, False)
    s = sw.Session(g, algo, mode=mode, budget=64 << 10, early_readback=1)
    v, st = s.run(src)
    s.close()
    assert np.array_equal(v, exp)
    if st.values_patched > 1:
        assert st.early_readback == 2
    assert st.early_readback in (1, 2)


def test_default_copies_small_graphs_at_the_end(sw, gpu):
    g, src, exp = graph(sw, "sssp", 20, 4, False)
    s = sw.Session(g, "sssp", mode="async", budget=4 << 20)
    v, st = s.run(src)
    s.close()
    assert st.early_readback == 0 and np.array_equal(v, exp)


@pytest.mark.parametrize("algo,mode", [("sssp", "async"), ("bfs", "sync"), ("cc", "async")])
@pytest.mark.parametrize("renumber", [False, True])
def test_default_on_a_graph_of_2_26_vertices(sw, gpu, algo, mode, renumber):
    """The default: a graph of 2^26 vertices copies its results early once
    the active edges fall to 1/64 of their peak (the async apps: after the
    first iteration; sync BFS: once its frontier has shrunk); two runs give
    the values of a session that copies at the end (whose answers the tests
    above pin to the oracle's)."""
    g = sw.Graph.rmat(26, 2, seed=7, weighted=algo == "sssp", symmetrize=algo == "cc")
    src = g.max_degree_vertex if algo != "cc" else 0
    if renumber:
        g.reorder()
    s0 = sw.Session(g, algo, mode=mode, budget=128 << 20, early_readback=-1)
    v0, st0 = s0.run(src)
    v0 = np.array(v0, copy=True)
    s0.close()
    assert st0.early_readback == 0
    s = sw.Session(g, algo, mode=mode, budget=128 << 20)
    for _ in range(2):
        v, st = s.run(src)
        assert np.array_equal(v, v0)
        if algo == "sssp":
            assert st.early_readback in (1, 2) and st.global_iterations > 1
    s.close()
