"""GPU parity of the schedule variants the autotuner and the learned policy
can pick, of the host-thread gather at the size where the reference drops a
tail, and of the result views a session hands out.  Everything is compared
with the CPU oracle (bit-exact for the label algorithms)."""
import gc

import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("unroll", [4, 8, 16, 32])
@pytest.mark.parametrize("budget", [0, 32 << 10])
def test_batch_every_unroll_matches_oracle(sw, gpu, unroll, budget):
    """Every wave assignment the autotuner may choose (4 and 8 positions per
    step run the load/decode phases, 16 and 32 the per-position path) gives
    the oracle's rows; the 32 KiB budget streams the hub vertex in chunks."""
    g = sw.Graph.rmat(13, 16, seed=23, weighted=True)
    ptr, adj, w = g.csr()
    # with 8-byte edges a buffer of the 32 KiB budget holds 2048 edges, fewer
    # than the hub has
    assert g.max_degree > (32 << 10) // 2 // 8
    rng = np.random.default_rng(unroll)
    cand = np.nonzero(np.diff(ptr) > 0)[0]
    sources = rng.choice(cand, 64, replace=False).astype(np.uint32)
    sources[0] = g.max_degree_vertex
    bs = sw.BatchSession(g, "sssp", budget=budget, unroll=unroll, packed=not budget)
    rows, st = bs.run(sources)
    bs.close()
    for i, s in enumerate(sources):
        exp = O.sssp(ptr, adj, w, int(s))
        assert np.array_equal(rows[i], exp), (unroll, budget, i, int(np.sum(rows[i] != exp)))
    if budget:
        assert st.partitions > st.global_iterations


def test_batch_tuned_variant_matches_oracle(sw, gpu):
    """After the autotuner has chosen, the session's runs use its choice and
    stay exact."""
    g = sw.Graph.rmat(12, 16, seed=8, weighted=True)
    ptr, adj, w = g.csr()
    sources = (np.arange(64, dtype=np.uint32) * 37) % g.n
    bs = sw.BatchSession(g, "bfs", budget=32 << 10)
    best, times = bs.tune(sources)
    assert best in (4, 8, 16, 32)
    rows, _ = bs.run(sources)
    bs.close()
    for i in range(64):
        assert np.array_equal(rows[i], O.bfs(ptr, adj, int(sources[i]))), i


def _overlap_runs(sw, algo, mode):
    """Learned schedules with fixed thresholds and any active volume allowed
    in place: whole id ranges move in units of one buffer, and later
    iterations group different ranges, so the two buffers can end up holding
    overlapping slices."""
    found = []
    for seed in range(6):
        for theta in (0.0, 0.2, 0.6):
            weighted = algo in ("sssp", "sswp")
            g = sw.Graph.rmat(13, 16, seed=100 + seed, weighted=weighted)
            ptr, adj, w = g.csr()
            src = g.max_degree_vertex
            s = sw.Session(g, algo, mode=mode, budget=40 << 10, policy="learned", theta=theta,
                           resident_min_bytes=1)
            vals, st = s.run(src)
            s.close()
            assert np.array_equal(vals, O.expected(algo, ptr, adj, w, src)), (seed, theta)
            found.append((st.resident_visits, st.resident_overlaps))
    return found


@pytest.mark.parametrize("algo,mode", [("sssp", "async"), ("bfs", "sync")])
def test_resident_visits_exact(sw, gpu, algo, mode):
    """In-place visits of slices a buffer still holds, admitted at any active
    volume, leave every result equal to the oracle's (the choice of which
    overlapping slice is visited is checked on the host by
    test_schedule_plan.py)."""
    found = _overlap_runs(sw, algo, mode)
    assert sum(v for v, _ in found) > 0, found           # in-place visits happened


def test_host_gather_large_unit(sw, gpu):
    """The host-thread gather with a unit of more than 50,000 active
    vertices whose count is not a multiple of 64 (the regime where the
    reference's chunking drops the last active vertices' edges,
    shared/subgraph_generator.cu:75-78,173-176): every vertex's value equals
    the oracle's.  The graph is a random relabelling of vertex ids, so the
    tail of the active list holds vertices with edges."""
    rng = np.random.default_rng(77)
    n, m = 300_007, 2_400_000
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    wts = rng.integers(1, 65, m).astype(np.uint32)
    g = sw.Graph.from_edges(n, src, dst, wts)
    ptr, adj, w = g.csr()
    s = sw.Session(g, "sssp", mode="sync", budget=64 << 20, transfer="host", resident=False)
    vals, st = s.run(0)
    s.close()
    assert np.array_equal(vals, O.sssp(ptr, adj, w, 0))
    big = [r["active_nodes"] for r in st.trace if r["active_nodes"] >= 50_000 and r["partitions"] == 1]
    assert big and any(a % 64 for a in big), [r["active_nodes"] for r in st.trace]
    assert st.h2d_bytes > 0


def test_result_view_outlives_session_reference(sw, gpu):
    """run(copy=False) returns a view of the session's host result array;
    the view keeps the session alive after the caller's last reference to it
    is gone, and still holds that run's values."""
    g = sw.Graph.rmat(12, 16, seed=4, weighted=True)
    ptr, adj, w = g.csr()
    src = g.max_degree_vertex
    view, _ = sw.Session(g, "sssp", mode="async", budget=64 << 10).run(src, copy=False)
    gc.collect()
    junk = [np.ones(1 << 20, np.uint32) for _ in range(8)]      # reuse freed host memory if any
    assert np.array_equal(view, O.sssp(ptr, adj, w, src))
    del junk
    assert not view.flags.writeable
    s = sw.Session(g, "bfs", mode="sync")
    s.close()
    with pytest.raises(sw.SubwayError):
        s.run(src)
