"""Learned scheduler on the GPU: the features the policy kernel records, the
fused inference against the stand-alone MLP kernels and the FP64 oracle, the
misprediction judge and its plain fallback, and bit-exact results under any
schedule the network picks."""
import os
import sys

import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import mlp_oracle as M  # noqa: E402


def _rmat(sw, scale, seed, weighted=True):
    return sw.Graph.rmat(scale, 16, seed=seed, weighted=weighted)


def test_features_of_first_iteration(sw, gpu):
    """Async runs activate every vertex for the first generate: the active
    counts, edge fraction and degree histogram are those of the whole graph."""
    g = _rmat(sw, 14, 3)
    ptr, adj, w = g.csr()
    deg = np.diff(ptr).astype(np.int64)
    vals, st = sw.run(g, "sssp", mode="async", source=g.max_degree_vertex, budget=128 << 10, policy="learned",
                      theta=2.0)
    assert np.array_equal(vals, O.sssp(ptr, adj, w, g.max_degree_vertex))
    f = np.array(st.trace[0]["features"], np.float64)
    a, ea = g.n, int(deg.sum())                    # every vertex is labelled, edges or not
    assert st.trace[0]["active_nodes"] == a and st.trace[0]["active_edges"] == ea
    assert f[1] == 0.0                             # not PageRank
    np.testing.assert_allclose(f[2], ea / g.m, rtol=1e-6)
    np.testing.assert_allclose(f[0], np.log2(1 + a) / np.log2(g.n), rtol=1e-5)
    d = deg
    hist = np.array([(d < 4).sum(), ((d >= 4) & (d < 32)).sum(), ((d >= 32) & (d < 256)).sum(), (d >= 256).sum()])
    np.testing.assert_allclose(f[4:8], hist / hist.sum(), rtol=1e-5)
    np.testing.assert_allclose(f[9], 1 / 16, rtol=1e-6)
    # later iterations see only their labelled vertices
    for r in st.trace[1:]:
        np.testing.assert_allclose(r["features"][2], r["active_edges"] / g.m, rtol=1e-5)


def test_fused_inference_matches_mlp_kernel_and_oracle(sw, gpu):
    """The heads the policy kernel computes from an iteration's features
    equal the stand-alone forward kernel's and the FP64 oracle's."""
    from subway_amd import policy
    g = _rmat(sw, 14, 5)
    rng = np.random.default_rng(2)
    p = policy.init_params(7, rng.random(12).astype(np.float32), 0.5 + rng.random(12).astype(np.float32))
    _, st = sw.run(g, "sssp", mode="async", source=g.max_degree_vertex, budget=128 << 10, policy="learned",
                   params=p)
    # iterations the network scheduled (after a fallback the plain policy runs
    # and no inference is made)
    used = [r for r in st.trace if not r["on_plain"]]
    assert len(used) >= 2
    x = np.array([r["features"] for r in used], np.float32)
    h = np.array([r["heads"] for r in used], np.float32)
    np.testing.assert_allclose(h, policy.forward(p, x), rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(h, M.forward(p.astype(np.float64), x.astype(np.float64))[0], rtol=1e-3, atol=1e-4)


@pytest.mark.parametrize("algo", ["sssp", "bfs", "sswp"])
def test_any_network_keeps_results_exact(sw, gpu, algo):
    """Random networks choose arbitrary thresholds and round caps; values
    stay bit-exact, and the decisions applied are the trained schedules
    nearest to the heads (theta in {0.25, 0.5, 0.8, 2}, cap in {1, 8, 64,
    4096}, nearest in log2)."""
    from subway_amd import policy
    weighted = algo != "bfs"
    g = _rmat(sw, 15, 9, weighted)
    ptr, adj, w = g.csr()
    src = g.max_degree_vertex
    exp = O.expected(algo, ptr, adj, w, src)
    for seed in range(3):
        p = policy.init_params(seed)
        vals, st = sw.run(g, algo, mode="async", source=src, budget=256 << 10, policy="learned", params=p)
        assert np.array_equal(vals, exp), (algo, seed)
        for r in st.trace:
            if r["on_plain"]:
                assert r["theta"] == 2.0
                continue
            thetas = np.array([0.25, 0.5, 0.8, 2.0])
            assert r["theta"] == pytest.approx(thetas[np.argmin(np.abs(thetas - r["heads"][1]))], rel=1e-6)
            caps = np.array([1, 8, 64, 4096])
            l2 = 16 * min(max(r["heads"][2], 0.0), 1.0)
            assert r["rounds_cap"] == caps[np.argmin(np.abs(np.log2(caps) - l2))]


def test_mispredictions_fall_back_to_plain(sw, gpu):
    """A network that always predicts an empty next frontier misses every
    prediction while PR-async (1e-6 threshold) keeps much of the graph
    active; once at least eight predictions have been judged and more than
    half of them missed, the run goes to the plain schedule for good."""
    from subway_amd import policy
    g = _rmat(sw, 14, 4, weighted=False)
    p = policy.init_params(1)
    p[policy.NORM:] = 0.0
    p[-3] = -1.0         # log2(next active-edge fraction) / 32: always 2^-32, nothing active
    p[-2] = 0.5          # theta 0.5
    p[-1] = 0.375        # cap 2^6
    ptr, adj, _ = g.csr()
    vals, st = sw.run(g, "pr", mode="async", budget=128 << 10, threshold=1e-6, policy="learned", params=p)
    exp, _ = O.pr(ptr, adj, 1e-6)
    assert (np.abs(vals.astype(np.float64) - exp) / exp).max() < 1e-4
    on = [r["on_plain"] for r in st.trace]
    assert 1 in on, (st.mispredictions, len(on))
    k = on.index(1)
    assert all(on[k:]) and not any(on[:k])
    assert st.fallback_iteration == k + 1 and k >= 8
    # the k predictions judged before the fallback: more than half missed
    assert 2 * st.mispredictions > k


def test_trained_network_is_shipped_and_exact(sw, gpu):
    """The committed scheduler network loads and keeps results exact."""
    from subway_amd import policy
    p = policy.default_params()
    if p is None:
        pytest.skip("no trained scheduler network in the package")
    g = _rmat(sw, 16, 12)
    ptr, adj, w = g.csr()
    src = g.max_degree_vertex
    vals, _ = sw.run(g, "sssp", mode="async", source=src, budget=512 << 10, policy="learned", params=p)
    assert np.array_equal(vals, O.sssp(ptr, adj, w, src))


def _median_ms(sess, src, reps=3):
    sess.run(src)
    t = []
    for _ in range(reps):
        v, st = sess.run(src)
        t.append(st.ms_total)
    return float(np.median(t)), v


def test_trained_network_gains_over_plain(sw, gpu):
    """What the shipped network gains over the plain policy, on graphs and
    seeds outside its training set: PR-async (whose plain schedule runs every
    partition to local convergence) finishes faster under the network's
    inner-round caps, within the 1e-4 PR bound of the FP64 oracle; SSSP-async
    is not slower.  Plain and learned sessions run interleaved."""
    from subway_amd import policy
    p = policy.default_params()
    if p is None:
        pytest.skip("no trained scheduler network in the package")
    g = sw.Graph.rmat(15, 16, seed=21)
    budget = g.m * 4
    plain = sw.Session(g, "pr", mode="async", budget=budget, threshold=1e-6)
    learned = sw.Session(g, "pr", mode="async", budget=budget, threshold=1e-6, policy="learned", params=p)
    tp, vp = _median_ms(plain, 0)
    tl, vl = _median_ms(learned, 0)
    ptr, adj, _ = g.csr()
    exp, _ = O.pr(ptr, adj, 1e-6)
    for v in (vp, vl):
        assert (np.abs(v.astype(np.float64) - exp) / exp).max() < 1e-4
    assert tp / tl > 1.1, (tp, tl)
    plain.close()
    learned.close()

    g = _rmat(sw, 20, 7)
    src = g.max_degree_vertex
    plain = sw.Session(g, "sssp", mode="async", budget=g.m * 2)
    learned = sw.Session(g, "sssp", mode="async", budget=g.m * 2, policy="learned", params=p)
    tp, vp = _median_ms(plain, src, 5)
    tl, vl = _median_ms(learned, src, 5)
    assert np.array_equal(vp, vl)
    assert tl <= 1.05 * tp, (tp, tl)


@pytest.mark.parametrize("algo,mode", [("sssp", "async"), ("bfs", "async"), ("cc", "async"), ("pr", "async"),
                                       ("sswp", "sync"), ("bfs", "sync")])
def test_learned_schedule_reads_resident_ranges_in_place(sw, gpu, algo, mode):
    """Under the learned schedule a range whose whole edge slice a buffer
    still holds from an earlier unit of the run is visited in place.  When
    the edges fit the two buffers, an async run transfers in its first
    iteration only (every vertex is active there, so the ranges move whole)
    when any active edge makes an in-place visit worth it (1 byte here);
    values equal the plain schedule's (PR: within the oracle bound); and
    resident=False moves the later iterations' edges again (the slice cache
    alone still skips a unit identical to the buffer's last one)."""
    g = _rmat(sw, 14, 9, weighted=algo in ("sssp", "sswp"))
    src = g.max_degree_vertex
    kw = dict(mode=mode, budget=int(g.m * 4 * 1.3), threshold=1e-6 if algo == "pr" else 0.01)
    vp, _ = sw.Session(g, algo, **kw).run(src)
    vl, st = sw.Session(g, algo, policy="learned", resident_min_bytes=1, **kw).run(src)
    vo, so = sw.Session(g, algo, policy="learned", resident=False, **kw).run(src)
    ptr, adj, w = g.csr()
    if algo == "pr":
        exp, _ = O.pr(ptr, adj, 1e-6)
        for v in (vp, vl, vo):
            assert (np.abs(v.astype(np.float64) - exp) / exp).max() < 1e-4
    else:
        exp = O.expected(algo, ptr, adj, w, src)
        for v in (vp, vl, vo):
            assert np.array_equal(v, exp)
    later = sum(r["h2d_bytes"] for r in st.trace[1:])
    later_off = sum(r["h2d_bytes"] for r in so.trace[1:])
    assert later <= later_off
    if mode == "async":
        assert st.global_iterations >= 2 and so.global_iterations >= 2
        assert st.trace[0]["h2d_bytes"] == g.m * st.edge_bytes
        assert later == 0
        if algo in ("sssp", "bfs"):
            # their later frontiers are sparse: without reuse they are gathered again
            assert later_off > 0


@pytest.mark.parametrize("buffers", [1, 2])
def test_resident_ranges_with_transfers_in_the_same_iteration(sw, gpu, buffers):
    """Resident ranges visited in place beside units that still transfer
    (the graph does not fit the buffers): with one buffer the in-place visit
    comes first and the transfers then reuse that buffer, with two the unit
    -> buffer map may rotate; values are the oracle's and the later
    iterations move no more than without the reuse."""
    g = _rmat(sw, 15, 4)
    src = g.max_degree_vertex
    kw = dict(mode="async", budget=int(g.m * 4 * 0.7), buffers=buffers, policy="learned")
    vl, st = sw.Session(g, "sssp", resident_min_bytes=1, **kw).run(src)
    vo, so = sw.Session(g, "sssp", resident=False, **kw).run(src)
    ptr, adj, w = g.csr()
    exp = O.sssp(ptr, adj, w, src)
    assert np.array_equal(vl, exp) and np.array_equal(vo, exp)
    assert sum(r["h2d_bytes"] for r in st.trace[1:]) <= sum(r["h2d_bytes"] for r in so.trace[1:])


def test_batched_learned_schedule_with_resident_ranges(sw, gpu):
    """Batched runs under the learned schedule visit resident ranges in place
    too (the persistent batched kernel over a whole id range): every source's
    row is the oracle's."""
    g = _rmat(sw, 14, 12)
    ptr, adj, w = g.csr()
    rng = np.random.default_rng(3)
    sources = rng.choice(np.nonzero(np.diff(ptr) > 0)[0], 16).astype(np.uint32)
    bs = sw.BatchSession(g, "sssp", budget=int(g.m * 4 * 1.3), policy="learned", resident_min_bytes=1)
    rows, st = bs.run(sources)
    for i, s in enumerate(sources):
        assert np.array_equal(rows[i], O.sssp(ptr, adj, w, int(s))), i
    assert st.global_iterations >= 2
    assert sum(r["h2d_bytes"] for r in st.trace[1:]) == 0
