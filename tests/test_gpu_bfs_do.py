"""Direction-optimizing BFS on a resident graph: top-down push steps and
bottom-up pull steps chosen by frontier size give the levels of the queue
BFS oracle, equal to the push-only levels, on R-MAT (configuration 1: scale
16, edge factor 16, source 0), random, disconnected and weighted graphs, in
both BFS modes."""
import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu


def _levels(sw, g, src, mode="sync", **kw):
    s = sw.Session(g, "bfs", mode=mode, **kw)
    best = None
    for _ in range(3):
        v, st = s.run(src)
        best = st if best is None or st.ms_total < best.ms_total else best
    s.close()
    return v, best


@pytest.mark.parametrize("mode", ["sync", "async"])
def test_config1_direction_optimizing(sw, gpu, mode):
    g = sw.Graph.rmat(16, 16, seed=42)
    ptr, adj, _ = g.csr()
    exp = O.bfs(ptr, adj, 0)
    do, st = _levels(sw, g, 0, mode)
    push, sp = _levels(sw, g, 0, mode, push_only=True)
    assert np.array_equal(do, exp) and np.array_equal(push, exp)
    assert st.pull_rounds >= 1 and sp.pull_rounds == 0
    assert st.pull_found > 0 and st.global_iterations == 1
    print(f"c1 bfs-{mode}: direction-optimizing {st.ms_total:.3f} ms ({st.pull_rounds} bottom-up steps, "
          f"{st.inner_rounds} levels), push only {sp.ms_total:.3f} ms")


@pytest.mark.parametrize("seed", range(4))
def test_random_and_disconnected(sw, gpu, seed):
    rng = np.random.default_rng(seed)
    n = 50_000
    m = int(n * (2 + 6 * seed))
    src = rng.integers(0, n // 2, m).astype(np.uint32)       # the upper half of the ids has no out-edges
    dst = rng.integers(0, n, m).astype(np.uint32)
    g = sw.Graph.from_edges(n, src, dst)
    ptr, adj, _ = g.csr()
    for s in (0, int(src[0]), n - 1):
        v, _ = _levels(sw, g, s)
        assert np.array_equal(v, O.bfs(ptr, adj, s)), (seed, s)


def test_weighted_graph_levels(sw, gpu):
    """BFS over a weighted graph (the resident buffer holds packed edges)."""
    g = sw.Graph.rmat(15, 16, seed=9, weighted=True)
    ptr, adj, _ = g.csr()
    src = g.max_degree_vertex
    v, st = _levels(sw, g, src, "async")
    assert np.array_equal(v, O.bfs(ptr, adj, src)) and st.pull_rounds >= 1


def test_partitioned_graph_unchanged(sw, gpu):
    """Graphs that do not fit one buffer keep the partitioned push path."""
    g = sw.Graph.rmat(14, 16, seed=3)
    ptr, adj, _ = g.csr()
    v, st = _levels(sw, g, 0, budget=64 << 10)
    assert np.array_equal(v, O.bfs(ptr, adj, 0)) and st.pull_rounds == 0
