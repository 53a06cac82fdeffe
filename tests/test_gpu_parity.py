"""GPU parity: every stage and every application against the CPU oracle,
through the C ABI.  Integer results are bit-exact; PageRank (FP32) stays
within 1e-4 max relative error of the FP64 oracle at the same threshold."""
import json
import os
import zlib

import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
PR_RTOL = 1e-4


def hand():
    with open(os.path.join(HERE, "golden", "hand_graphs.json")) as f:
        return json.load(f)


def graph_of(sw, n, edges, weighted):
    src = np.array([e[0] for e in edges], np.uint32)
    dst = np.array([e[1] for e in edges], np.uint32)
    w = np.array([e[2] for e in edges], np.uint32) if weighted else None
    return sw.Graph.from_edges(n, src, dst, w)


def random_graph(sw, rng, n, m, weighted=True, hub=None):
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    if hub is not None:
        v, d = hub
        src = np.concatenate([src, np.full(d, v, np.uint32)])
        dst = np.concatenate([dst, rng.integers(0, n, d).astype(np.uint32)])
    w = rng.integers(1, 65, len(src)).astype(np.uint32) if weighted else None
    return sw.Graph.from_edges(n, src, dst, w)


def check_app(sw, g, algo, mode, src=0, **kw):
    vals, st = sw.run(g, algo, mode=mode, source=src, **kw)
    ptr, adj, w = g.csr()
    if algo == "pr":
        acc = kw.get("threshold", 0.01)
        exp, _ = O.pr(ptr, adj, acc)
        rel = np.abs(vals.astype(np.float64) - exp) / np.maximum(exp, 1e-30)
        assert rel.max() < PR_RTOL, (algo, mode, float(rel.max()))
    else:
        exp = O.expected(algo, ptr, adj, w, src)
        bad = np.nonzero(vals != exp)[0]
        assert bad.size == 0, (algo, mode, kw, bad[:10], vals[bad[:10]], exp[bad[:10]])
    return vals, st


# ------------------------------------------------------------- generator
@pytest.mark.parametrize("n", [1, 15, 16, 17, 4095, 4096, 4097, 100_000, 1_048_583])
@pytest.mark.parametrize("density", [0.0, 0.003, 0.5, 1.0])
def test_generate_matches_oracle(sw, gpu, n, density):
    rng = np.random.default_rng(n + int(density * 1000))
    l1 = (rng.random(n) < density / 2).astype(np.uint8)
    l2 = (rng.random(n) < density / 2).astype(np.uint8)
    if density == 1.0:
        l1[:] = 1
    deg = rng.integers(0, 1000, n).astype(np.uint32)
    an, ap = sw.generate(l1, l2, deg)
    ean, eap = O.generate(l1 | l2, deg)
    assert (an == ean).all() and (ap == eap).all()


def test_generate_pr_matches_oracle(sw, gpu):
    rng = np.random.default_rng(5)
    n = 300_001
    delta = rng.random(n).astype(np.float32) * 0.02
    deg = rng.integers(0, 30, n).astype(np.uint32)
    an, ap = sw.generate_pr(delta, 0.01, deg)
    ean, eap = O.generate((delta > np.float32(0.01)).astype(np.uint8), deg)
    assert (an == ean).all() and (ap == eap).all()


def test_generate_64bit_offsets(sw, gpu):
    """Active edge totals past 2^32 (the reference's u32 scan overflows)."""
    n = 3_000_000
    deg = np.full(n, 3000, np.uint32)
    ones = np.ones(n, np.uint8)
    an, ap = sw.generate(ones, np.zeros(n, np.uint8), deg)
    assert len(an) == n and int(ap[-1]) == 3000 * n > 2**32
    assert (ap[1:] - ap[:-1] == 3000).all()


def test_generate_repeated_launches(sw, gpu):
    """Epoch-tagged descriptors: many launches in a row stay exact."""
    rng = np.random.default_rng(9)
    n = 50_000
    deg = rng.integers(0, 10, n).astype(np.uint32)
    for i in range(20):
        l1 = (rng.random(n) < 0.3).astype(np.uint8)
        an, ap = sw.generate(l1, np.zeros(n, np.uint8), deg)
        ean, eap = O.generate(l1, deg)
        assert (an == ean).all() and (ap == eap).all()


# ----------------------------------------------------------- partitioner
@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("budget", [1, 64, 5000, 10**12])
def test_partition_matches_oracle(sw, gpu, seed, budget):
    rng = np.random.default_rng(seed)
    a = int(rng.integers(1, 200_000))
    deg = rng.integers(0, 300, a).astype(np.uint64)
    deg[rng.integers(0, a, 3)] = 20_000
    ap = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    an = np.sort(rng.choice(10 * a, a, replace=False)).astype(np.uint32)
    rows = sw.partition(ap, an, budget)
    exp = O.partition(ap, budget)
    assert rows.shape[0] == exp.shape[0]
    assert (rows[:, :4] == exp).all()
    assert (rows[:, 4] == an[rows[:, 0].astype(np.int64)]).all()
    assert (rows[:, 5] == an[(rows[:, 0] + rows[:, 1] - 1).astype(np.int64)]).all()


# ---------------------------------------------------------------- gather
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("frac", [0.0005, 0.05, 0.6, 1.0])
def test_gather_matches_csr_slices(sw, gpu, weighted, frac):
    g = sw.Graph.rmat(14, 16, seed=3, weighted=weighted)
    ptr, adj, w = g.csr()
    rng = np.random.default_rng(int(frac * 1e4))
    act = np.nonzero(rng.random(g.n) < frac)[0].astype(np.uint32)
    if act.size == 0:
        act = np.array([0], np.uint32)
    ends, wts = sw.gather(g, act)
    idx = np.concatenate([np.arange(ptr[v], ptr[v + 1]) for v in act]).astype(np.int64)
    assert (ends == adj[idx]).all()
    if weighted:
        assert (wts == w[idx]).all()


# ---------------------------------------------------------- applications
APPS = [(a, m) for a in ("bfs", "cc", "sssp", "sswp") for m in ("sync", "async")]


@pytest.mark.parametrize("name", ["readme", "detour", "cycles", "wide"])
@pytest.mark.parametrize("algo,mode", APPS)
def test_apps_golden(sw, gpu, name, algo, mode):
    gd = hand()[name]
    g = graph_of(sw, gd["n"], gd["edges"], weighted=True)
    vals, _ = sw.run(g, algo, mode=mode, source=gd["source"])
    assert vals.tolist() == gd[algo]


EDGE_CASES = {
    # name: (n, edges (src, dst, w), source)
    "single_vertex": (1, [], 0),
    "no_edges": (5, [], 3),
    "isolated_source": (6, [(0, 1, 3), (1, 2, 4), (2, 0, 1), (4, 5, 9)], 3),
    "self_loops_only": (3, [(0, 0, 5), (1, 1, 2)], 0),
    "sink_source": (4, [(0, 1, 1), (1, 2, 1), (2, 3, 1)], 3),
}


@pytest.mark.parametrize("case", sorted(EDGE_CASES))
@pytest.mark.parametrize("algo,mode", APPS + [("pr", "sync"), ("pr", "async")])
def test_apps_edge_cases(sw, gpu, case, algo, mode):
    """Degenerate inputs: one vertex, no edges at all, a source without
    edges, only self-loops, a source that is a sink; every application
    matches the oracle (PR to its tolerance)."""
    n, edges, source = EDGE_CASES[case]
    src = np.array([e[0] for e in edges], np.uint32)
    dst = np.array([e[1] for e in edges], np.uint32)
    w = np.array([e[2] for e in edges], np.uint32) if algo in ("sssp", "sswp") else None
    g = sw.Graph.from_edges(n, src, dst, w)
    kw = {"threshold": 1e-6} if algo == "pr" else {}
    check_app(sw, g, algo, mode, src=source if algo != "cc" else 0, **kw)


@pytest.mark.parametrize("mode", ["sync", "async"])
def test_pr_golden_readme(sw, gpu, mode):
    gd = hand()["readme"]
    g = graph_of(sw, gd["n"], gd["edges"], weighted=False)
    vals, _ = sw.run(g, "pr", mode=mode, threshold=1e-9)
    np.testing.assert_allclose(vals, gd["pr"], rtol=1e-6)


@pytest.mark.parametrize("algo,mode", APPS + [("pr", "sync"), ("pr", "async")])
@pytest.mark.parametrize("budget", [0, 64 << 10])
@pytest.mark.parametrize("transfer", ["auto", "zerocopy", "host"])
def test_apps_random(sw, gpu, algo, mode, budget, transfer):
    rng = np.random.default_rng(zlib.crc32(f"{algo}-{mode}-{budget}-{transfer}".encode()))
    n = 20_000
    weighted = algo in ("sssp", "sswp")
    g = random_graph(sw, rng, n, 8 * n, weighted=weighted, hub=(17, 9000))
    kw = dict(budget=budget, transfer=transfer)
    if algo == "pr":
        kw["threshold"] = 1e-6
    _, st = check_app(sw, g, algo, mode, src=17 if algo != "cc" else 0, **kw)
    if budget:
        assert st.partitions > st.global_iterations  # the out-of-memory path ran


ASYNC = [(a, "async") for a in ("bfs", "cc", "sssp", "sswp", "pr")]


@pytest.mark.parametrize("algo,mode", APPS + [("pr", "sync"), ("pr", "async")])
def test_resident_graph(sw, gpu, algo, mode):
    """A graph whose edges fit one device buffer moves once per run and
    stays resident; every partition visit then runs rounds to the fixpoint
    on it.  Values equal the oracle's and the per-iteration transfers'."""
    weighted = algo in ("sssp", "sswp")
    g = sw.Graph.rmat(14, 16, seed=41, weighted=weighted, symmetrize=(algo == "cc"))
    src = g.max_degree_vertex if algo != "cc" else 0
    kw = {"threshold": 1e-6} if algo == "pr" else {}
    s = sw.Session(g, algo, mode=mode, budget=8 << 20, **kw)
    for _ in range(2):                       # the buffer is loaded again in every run
        v, st = s.run(src)
        assert st.h2d_bytes == st.edge_bytes * g.m
        ptr, adj, w = g.csr()
        if algo == "pr":
            exp, _ = O.pr(ptr, adj, 1e-6)
            assert (np.abs(v.astype(np.float64) - exp) / exp).max() < PR_RTOL
        else:
            assert np.array_equal(v, O.expected(algo, ptr, adj, w, src))
    s.close()
    vn, sn = check_app(sw, g, algo, mode, src=src, budget=8 << 20, resident=False, **kw)
    if algo != "pr":
        assert np.array_equal(v, vn)


@pytest.mark.parametrize("algo,mode", ASYNC)
@pytest.mark.parametrize("rounds", ["persistent", "per-round"])
@pytest.mark.parametrize("budget", [0, 48 << 10])
def test_async_rounds_modes(sw, gpu, algo, mode, rounds, budget):
    """The persistent rounds kernel and the launch-per-round
    path reach the same fixpoint as the oracle."""
    weighted = algo in ("sssp", "sswp")
    g = sw.Graph.rmat(13, 16, seed=21, weighted=weighted, symmetrize=(algo == "cc"))
    kw = dict(budget=budget, rounds=rounds)
    if algo == "pr":
        kw["threshold"] = 1e-6
    check_app(sw, g, algo, mode, src=g.max_degree_vertex, **kw)


@pytest.mark.parametrize("max_rounds", [1, 2, 3])
def test_async_round_cap_still_exact(sw, gpu, max_rounds):
    """Capping the inner rounds per partition only changes the schedule:
    labels left set are picked up by the next generate."""
    g = sw.Graph.rmat(13, 16, seed=4, weighted=True)
    _, st = check_app(sw, g, "sssp", "async", src=0, budget=64 << 10, max_rounds=max_rounds)
    _, full = sw.run(g, "sssp", mode="async", source=0, budget=64 << 10)
    assert st.global_iterations >= full.global_iterations


@pytest.mark.parametrize("algo", ["bfs", "sssp", "sswp"])
@pytest.mark.parametrize("mode", ["sync", "async"])
def test_session_many_sources(sw, gpu, algo, mode):
    """One session (device state built once) answers queries from several
    sources, each bit-exact."""
    weighted = algo in ("sssp", "sswp")
    g = sw.Graph.rmat(13, 16, seed=33, weighted=weighted)
    sess = sw.Session(g, algo, mode=mode, budget=64 << 10)
    ptr, adj, w = g.csr()
    rng = np.random.default_rng(1)
    for i, src in enumerate([g.max_degree_vertex] + rng.integers(0, g.n, 4).tolist()):
        vals, _ = sess.run(int(src), copy=bool(i % 2))   # odd runs: a view of the host result array
        assert np.array_equal(vals, O.expected(algo, ptr, adj, w, int(src)))


@pytest.mark.parametrize("algo,mode", APPS)
@pytest.mark.parametrize("buffers", [1, 2])
def test_apps_rmat_tiny_budget(sw, gpu, algo, mode, buffers):
    """R-MAT scale 14 through a 32 KiB budget: dozens of partitions and
    hub vertices larger than a buffer (streamed in chunks)."""
    weighted = algo in ("sssp", "sswp")
    g = sw.Graph.rmat(14, 16, seed=11, weighted=weighted, symmetrize=(algo == "cc"))
    assert g.max_degree > 2048
    check_app(sw, g, algo, mode, src=g.max_degree_vertex, budget=32 << 10, buffers=buffers)


def test_config1_bfs_sync_scale16(sw, gpu):
    """R-MAT scale 16, edge factor 16, source 0, no budget cap: bit-exact
    against the oracle and the OpenMP build.  The graph fits the device, so
    it stays resident and one visit runs the rounds to the fixpoint; with
    per-iteration transfers (the reference's loop) the same levels take
    several global iterations."""
    g = sw.Graph.rmat(16, 16, seed=42)
    vals, st = check_app(sw, g, "bfs", "sync", src=0)
    ptr, adj, _ = g.csr()
    omp, _, _ = O.omp_label("bfs", ptr, adj, None, 0)
    assert (vals == omp).all()
    assert st.global_iterations == 1 and st.inner_rounds >= 2
    vn, sn = check_app(sw, g, "bfs", "sync", src=0, resident=False)
    assert np.array_equal(vn, vals) and sn.global_iterations >= 2


def test_deterministic_repeats(sw, gpu):
    g = sw.Graph.rmat(15, 16, seed=5, weighted=True)
    first, _ = sw.run(g, "sssp", mode="async", budget=256 << 10)
    for _ in range(3):
        again, _ = sw.run(g, "sssp", mode="async", budget=256 << 10)
        assert (again == first).all()


def test_pr_rmat_tolerance(sw, gpu):
    g = sw.Graph.rmat(14, 16, seed=8)
    check_app(sw, g, "pr", "async", threshold=1e-6, budget=64 << 10)
    check_app(sw, g, "pr", "sync", threshold=1e-6)


@pytest.mark.parametrize("max_rounds", [4096, 8])
def test_pr_async_keeps_hub_mass(sw, gpu, max_rounds):
    """Async PR with partitions run to local convergence (thousands of tiny
    pushes into a hub that is not resident) keeps the rank mass: the error
    against the converged FP64 ranks and the missing mass stay those of the
    FP64 oracle at the same threshold (about 3e-6), far inside the 1e-4
    bound.  An FP32 residual rounds away shares below half the hub's ulp
    here and misses by 8e-5 at the hub."""
    g = sw.Graph.rmat(15, 16, seed=21)
    ptr, adj, _ = g.csr()
    exact, _ = O.pr(ptr, adj, 1e-10)
    at_tau, _ = O.pr(ptr, adj, 1e-6)
    v, _ = sw.run(g, "pr", mode="async", budget=g.m * 4, threshold=1e-6, max_rounds=max_rounds)
    v = v.astype(np.float64)
    own = np.abs(at_tau - exact) / exact
    err = np.abs(v - exact) / exact
    assert err.max() < 2 * own.max() + 1e-6, (err.max(), own.max())
    deficit = (exact.sum() - v.sum()) / exact.sum()
    oracle_deficit = (exact.sum() - at_tau.sum()) / exact.sum()
    assert abs(deficit) < 2 * abs(oracle_deficit) + 1e-6, (deficit, oracle_deficit)


@pytest.mark.parametrize("algo", ["sssp", "sswp", "bfs", "cc"])
@pytest.mark.parametrize("mode", ["sync", "async"])
def test_packed_edges_match_unpacked(sw, gpu, algo, mode):
    """Weighted edges moved and buffered as one word (destination |
    (weight - 1) << 14 for this 16384-vertex graph) give the oracle's values,
    and the first iteration (the whole graph) moves exactly half the bytes."""
    g = sw.Graph.rmat(14, 16, seed=31, weighted=True, symmetrize=(algo == "cc"))
    src = g.max_degree_vertex if algo != "cc" else 0
    vp, sp = check_app(sw, g, algo, mode, src=src, budget=256 << 10, packed=True)
    vu, su = check_app(sw, g, algo, mode, src=src, budget=256 << 10, packed=False)
    assert np.array_equal(vp, vu)
    if mode == "async":
        assert sp.trace[0]["h2d_bytes"] * 2 == su.trace[0]["h2d_bytes"] == 8 * g.m


@pytest.mark.parametrize("algo", ["sssp", "sswp"])
def test_packed_weight_span(sw, gpu, algo):
    """What has to fit beside the ids is the span of the weights, not the
    largest weight: ids of 20 bits leave 12, and weights 5000 .. 9095 pack
    (stored less 5000) though 9095 needs 14 bits."""
    rng = np.random.default_rng(12)
    n = 1 << 20
    src = rng.integers(0, 4096, 40_000).astype(np.uint32)
    dst = rng.integers(0, 8192, 40_000).astype(np.uint32)
    w = rng.integers(5000, 5000 + 4096, 40_000).astype(np.uint32)
    w[:2] = (5000, 5000 + 4095)
    g = sw.Graph.from_edges(n, src, dst, w)
    s0 = int(src[np.argmax(np.bincount(src))])
    vp, st = check_app(sw, g, algo, "async", src=s0, budget=64 << 10)
    assert st.trace[0]["h2d_bytes"] == 4 * g.m
    vu, _ = check_app(sw, g, algo, "async", src=s0, budget=64 << 10, packed=False)
    assert np.array_equal(vp, vu)


@pytest.mark.parametrize("transfer", ["auto", "zerocopy", "host"])
@pytest.mark.parametrize("algo,mode", [("sssp", "async"), ("sswp", "async"), ("sssp", "sync")])
def test_split_edges(sw, gpu, algo, mode, transfer):
    """Ids of 25 bits leave 7 for the weights and the weights 1 .. 255 need
    8: the edges do not pack, and move as a u32 destination array plus a u8
    weight array (5 bytes per edge), by DMA, GPU gather and host gather,
    including a vertex larger than a buffer; values equal the oracle's and
    the 8-byte format's."""
    rng = np.random.default_rng(31)
    n = 1 << 25
    m = 60_000
    src = np.concatenate([rng.integers(0, 4096, m), np.full(10_000, 7)]).astype(np.uint32)
    dst = np.concatenate([rng.integers(0, 8192, m), rng.integers(0, n, 10_000)]).astype(np.uint32)
    w = rng.integers(1, 256, len(src)).astype(np.uint32)
    w[:2] = (1, 255)
    g = sw.Graph.from_edges(n, src, dst, w)
    kw = dict(budget=64 << 10, transfer=transfer)
    vs, st = check_app(sw, g, algo, mode, src=7, **kw)
    assert st.edge_bytes == 5 and st.partitions > st.global_iterations
    if mode == "async":
        assert st.trace[0]["h2d_bytes"] == 5 * g.m
    vu, su = check_app(sw, g, algo, mode, src=7, packed=False, **kw)
    assert su.edge_bytes == 8 and np.array_equal(vs, vu)


@pytest.mark.parametrize("algo,mode", [("sssp", "async"), ("sswp", "async"), ("sssp", "sync"), ("bfs", "async")])
def test_split_layout_graph(sw, gpu, algo, mode):
    """A graph laid out for the split format only (pinned destination and
    weight arrays, the 8-byte list in ordinary memory — what large weighted
    graphs that do not pack get from the generator) gives the oracle's
    values in the split format and refuses 8-byte transfers."""
    g = sw.Graph.rmat(15, 16, seed=5, weighted=True, split_layout=True)
    src = g.max_degree_vertex
    _, st = check_app(sw, g, algo, mode, src=src, budget=96 << 10)
    assert st.edge_bytes == 5 and st.partitions > st.global_iterations
    with pytest.raises(sw.SubwayError, match="split format"):
        sw.run(g, algo, mode=mode, source=src, budget=96 << 10, packed=False)


def test_pair_counters_split_the_work(sw, gpu):
    """The work of stand-alone frontier + advance launches is counted apart
    from the persistent launches': all of it in launch-per-round mode and
    with a round cap of one, none of it in a persistent run without
    oversized vertices; the values are the oracle's in all three."""
    g = sw.Graph.rmat(14, 16, seed=3, weighted=True)
    src = g.max_degree_vertex
    kw = dict(src=src, budget=128 << 10)
    for extra in (dict(rounds="per-round"), dict(max_rounds=1)):
        _, st = check_app(sw, g, "sssp", "async", **kw, **extra)
        assert st.work_edges > 0
        assert (st.pair_work_edges, st.pair_work_entries, st.pair_updates) == \
            (st.work_edges, st.work_entries, st.updates), extra
    _, st = check_app(sw, g, "sssp", "async", **kw)
    assert st.work_edges > 0
    assert (st.pair_work_edges, st.pair_work_entries, st.pair_updates) == (0, 0, 0)


@pytest.mark.parametrize("ext", ["bwcsr", "bwcsr64"])
def test_split_layout_file(sw, gpu, tmp_path, ext):
    """A .bwcsr file read in the split layout (the layout a weighted file of
    16 GiB or more that does not pack gets by itself) gives the oracle's
    values with 5-byte transfers, the same as the file read as usual; weights
    beyond a byte are refused in that layout."""
    g = sw.Graph.rmat(15, 16, seed=5, weighted=True)
    path = str(tmp_path / f"g.{ext}")
    g.save(path)
    gs = sw.Graph.load(path, split_layout=True)
    src = g.max_degree_vertex
    vs, st = check_app(sw, gs, "sssp", "async", src=src, budget=96 << 10)
    assert st.edge_bytes == 5 and st.partitions > st.global_iterations
    vu, _ = check_app(sw, sw.Graph.load(path), "sssp", "async", src=src, budget=96 << 10)
    assert np.array_equal(vs, vu)
    with pytest.raises(sw.SubwayError, match="split format"):
        sw.run(gs, "sssp", mode="async", source=src, budget=96 << 10, packed=False)
    heavy = sw.Graph.rmat(10, 8, seed=6, weighted=True, wmax=300)
    hp = str(tmp_path / "heavy.bwcsr")
    heavy.save(hp)
    with pytest.raises(sw.SubwayError, match="fit a byte"):
        sw.Graph.load(hp, split_layout=True)


def test_unpackable_weights_fall_back(sw, gpu):
    """Weights that fit neither beside the ids nor in a byte keep the 8-byte
    format."""
    rng = np.random.default_rng(8)
    n = 1 << 20                               # ids need 20 bits, 12 left for weights
    src = rng.integers(0, 4096, 30_000).astype(np.uint32)
    dst = rng.integers(0, n, 30_000).astype(np.uint32)
    w = rng.integers(1, 1 << 16, 30_000).astype(np.uint32)
    g = sw.Graph.from_edges(n, src, dst, w)
    _, st = check_app(sw, g, "sssp", "async", src=int(src[0]), budget=64 << 10)
    assert st.trace[0]["h2d_bytes"] == 8 * g.m


def test_bounded_wait_reports_error(sw, gpu):
    """Every device-side wait (scan lookback, grid barrier) gives up after a
    poll limit: with a limit of 1 poll the grid barrier of the persistent
    rounds kernel (over a thousand workgroups) gives up, every kernel still
    runs to its end, the run fails with an error naming the wait, and a new
    run with the default limit is exact."""
    g = sw.Graph.rmat(15, 16, seed=5, weighted=True)
    with pytest.raises(sw.SubwayError, match="device wait gave up"):
        sw.run(g, "sssp", mode="async", source=g.max_degree_vertex, budget=512 << 10, spin_limit=1)
    check_app(sw, g, "sssp", "async", src=g.max_degree_vertex, budget=512 << 10)


@pytest.mark.parametrize("algo", ["sssp", "cc", "pr"])
def test_persistent_launch_with_few_members(sw, gpu, algo):
    """A persistent launch whose first workgroup closes membership after one
    tick of the 100 MHz clock runs with whatever part of the grid had
    started by then (as when kernels of other streams hold slots): the
    members claim all the work, every barrier counts only them, and the
    fixpoint is the oracle's."""
    weighted = algo == "sssp"
    g = sw.Graph.rmat(14, 16, seed=23, weighted=weighted, symmetrize=(algo == "cc"))
    kw = dict(budget=128 << 10, join_window=1)
    if algo == "pr":
        kw["threshold"] = 1e-6
    _, st = check_app(sw, g, algo, "async", src=g.max_degree_vertex if algo != "cc" else 0, **kw)
    _, full = check_app(sw, g, algo, "async", src=g.max_degree_vertex if algo != "cc" else 0, budget=128 << 10,
                        **({"threshold": 1e-6} if algo == "pr" else {}))
    assert 1 <= st.min_members < full.min_members


def test_batch_with_few_members(sw, gpu):
    g = sw.Graph.rmat(13, 16, seed=17, weighted=True)
    ptr, adj, w = g.csr()
    sources = np.arange(0, 64 * 97, 97, dtype=np.uint32) % g.n
    bs = sw.BatchSession(g, "sssp", budget=64 << 10, join_window=1)
    vals, st = bs.run(sources)
    bs.close()
    for i in (0, 17, 63):
        assert np.array_equal(vals[i], O.sssp(ptr, adj, w, int(sources[i])))
    assert st.min_members >= 1


def test_errors_are_reported(sw, gpu):
    g = sw.Graph.rmat(10, 4, seed=1)
    with pytest.raises(sw.SubwayError):
        sw.run(g, "sssp")  # unweighted graph
    with pytest.raises(sw.SubwayError):
        sw.run(g, "bfs", source=g.n)  # source out of range
    with pytest.raises(sw.SubwayError):
        sw.run(g, "bfs", budget=100)  # below one advance chunk per buffer


# ------------------------------------------------------ batched sources
@pytest.mark.parametrize("algo", ["sssp", "bfs", "sswp"])
@pytest.mark.parametrize("budget", [0, 64 << 10, 32 << 10])
@pytest.mark.parametrize("k", [1, 5, 64])
def test_batch_matches_single_source(sw, gpu, algo, budget, k):
    """k sources advanced together equal k single-source oracle runs; the
    32 KiB budget makes the hub vertex stream in chunks."""
    g = sw.Graph.rmat(13, 16, seed=17, weighted=True)
    ptr, adj, w = g.csr()
    deg = np.diff(ptr)
    rng = np.random.default_rng(k + budget)
    cand = np.nonzero(deg > 0)[0]
    sources = rng.choice(cand, k, replace=True).astype(np.uint32)
    sources[0] = g.max_degree_vertex
    bs = sw.BatchSession(g, algo, budget=budget)
    vals, st = bs.run(sources)
    for i, s in enumerate(sources):
        exp = O.expected(algo, ptr, adj, w, int(s))
        assert np.array_equal(vals[i], exp), (algo, budget, k, i, int(np.sum(vals[i] != exp)))
    # per-source work counters: one source bit per pushing vertex when k = 1,
    # at most k otherwise, and every successful source update was a relaxation
    assert st.work_entries <= st.lane_entries <= k * st.work_entries
    assert st.work_edges <= st.lane_work <= k * st.work_edges
    assert st.updates <= st.lane_work
    if k == 1:
        assert (st.lane_work, st.lane_entries) == (st.work_edges, st.work_entries)


def test_batch_graph_replays_match_eager(sw, gpu):
    """Batched iterations start with a recorded HIP graph (union generate,
    error words, partition table), replayed every iteration of every run;
    with profiling the same launches run eagerly.  Rows are identical, run
    after run, with eager runs of another session in between."""
    g = sw.Graph.rmat(13, 16, seed=29, weighted=True)
    sources = np.arange(0, 64 * 61, 61, dtype=np.uint32) % g.n
    graphed = sw.BatchSession(g, "sssp", budget=48 << 10)
    eager = sw.BatchSession(g, "sssp", budget=48 << 10, profile=True)
    ref, st0 = eager.run(sources)
    assert st0.global_iterations >= 3
    for i in range(4):
        rows, st = graphed.run(sources)
        assert np.array_equal(rows, ref), i
        again, _ = eager.run(sources)
        assert np.array_equal(again, ref), i
    graphed.close()
    eager.close()


def test_batch_tuner_and_repeats(sw, gpu):
    g = sw.Graph.rmat(12, 16, seed=2, weighted=True)
    ptr, adj, w = g.csr()
    sources = np.arange(0, 64 * 13, 13, dtype=np.uint32) % g.n
    bs = sw.BatchSession(g, "sssp", budget=128 << 10)
    best, times = bs.tune(sources)
    assert best in (4, 8, 16, 32) and all(t > 0 for t in times)
    for _ in range(2):
        vals, _ = bs.run(sources)
        for i in (0, 17, 63):
            assert np.array_equal(vals[i], O.sssp(ptr, adj, w, int(sources[i])))


# ------------------------------------------------- learned-policy schedule
@pytest.mark.parametrize("algo,mode", APPS + [("pr", "sync"), ("pr", "async")])
@pytest.mark.parametrize("theta", [0.0, 0.3, 1.5])
def test_range_schedule_exact(sw, gpu, algo, mode, theta):
    """The learned policy's schedule (static id ranges; dense ranges move
    whole and their rounds cover every vertex; the rest is compacted) only
    changes the schedule: results equal the oracle for any threshold."""
    weighted = algo in ("sssp", "sswp")
    g = sw.Graph.rmat(13, 16, seed=29, weighted=weighted, symmetrize=(algo == "cc"))
    kw = dict(budget=48 << 10, policy="learned", theta=theta)
    if algo == "pr":
        kw["threshold"] = 1e-6
    _, st = check_app(sw, g, algo, mode, src=g.max_degree_vertex, **kw)
    assert st.partitions >= st.global_iterations


@pytest.mark.parametrize("theta", [0.0, 0.4])
def test_range_schedule_batched(sw, gpu, theta):
    g = sw.Graph.rmat(12, 16, seed=31, weighted=True)
    ptr, adj, w = g.csr()
    sources = np.arange(0, 640, 10, dtype=np.uint32)
    bs = sw.BatchSession(g, "sssp", budget=48 << 10, policy="learned", theta=theta)
    vals, _ = bs.run(sources)
    for i in (0, 9, 63):
        assert np.array_equal(vals[i], O.sssp(ptr, adj, w, int(sources[i])))


# ------------------------------------------------------ application binaries
APP_BIN = os.path.join(os.path.dirname(HERE), "subway_amd", "bin")


@pytest.mark.parametrize("app", [f"{a}-{m}" for a in ("bfs", "cc", "sssp", "sswp", "pr") for m in ("sync", "async")])
def test_app_binaries(sw, gpu, tmp_path, app):
    """The ten applications as the reference's users run them: --input a
    converted binary graph, --source, --output; the output file holds one
    "<vertex> <value>" line per vertex (ostream formatting, 6 significant
    digits for PR) with the oracle's values."""
    import subprocess
    algo, mode = app.split("-")
    weighted = algo in ("sssp", "sswp")
    g = sw.Graph.rmat(12, 8, seed=3, weighted=weighted, symmetrize=(algo == "cc"))
    path = tmp_path / ("g.bwcsr" if weighted else "g.bcsr")
    g.save(str(path))
    out = tmp_path / "out.txt"
    src = g.max_degree_vertex if algo != "cc" else 0
    cmd = [os.path.join(APP_BIN, app), "--input", str(path), "--output", str(out), "--budget", "64K"]
    if not (algo == "cc" and mode == "sync"):
        cmd += ["--source", str(src)]
    if algo == "pr":
        cmd += ["--threshold", "1e-6"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Processing finished in" in r.stdout
    rows = [ln.split() for ln in out.read_text().splitlines()]
    assert len(rows) == g.n and all(int(v) == i for i, (v, _) in enumerate(rows))
    ptr, adj, w = g.csr()
    if algo == "pr":
        vals = np.array([float(x) for _, x in rows])
        exp, _ = O.pr(ptr, adj, 1e-6)
        assert (np.abs(vals - exp) / exp).max() < PR_RTOL + 1e-5     # plus the 6-digit printing
    else:
        vals = np.array([int(x) for _, x in rows], dtype=np.uint64)
        assert np.array_equal(vals, O.expected(algo, ptr, adj, w, src).astype(np.uint64))
