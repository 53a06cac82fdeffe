"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the parity
reference: hand-derived golden answers, networkx, brute force on tiny
graphs, a closed-form PageRank solve, and planted mistakes that must fail."""
import itertools
import json
import os

import numpy as np
import pytest

from tests import _oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
INF = O.INF


def load_hand():
    with open(os.path.join(HERE, "golden", "hand_graphs.json")) as f:
        return json.load(f)


def to_csr(n, edges):
    edges = sorted(edges, key=lambda e: e[0])  # stable: keeps in-source order
    ptr = np.zeros(n + 1, np.uint64)
    for e in edges:
        ptr[e[0] + 1] += 1
    ptr = np.cumsum(ptr).astype(np.uint64)
    adj = np.array([e[1] for e in edges], np.uint32)
    w = np.array([e[2] if len(e) > 2 else 1 for e in edges], np.uint32)
    return ptr, adj, w


def random_graph(rng, n, m, wmax=64):
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    w = rng.integers(1, wmax + 1, m)
    return to_csr(n, list(zip(src.tolist(), dst.tolist(), w.tolist())))


@pytest.mark.parametrize("name", ["readme", "detour", "cycles", "wide"])
@pytest.mark.parametrize("algo", ["bfs", "sssp", "sswp", "cc"])
def test_oracle_hand_graphs(name, algo):
    g = load_hand()[name]
    ptr, adj, w = to_csr(g["n"], g["edges"])
    got = O.expected(algo, ptr, adj, w, g["source"])
    assert got.tolist() == g[algo]


def test_oracle_pr_readme_closed_form():
    g = load_hand()["readme"]
    ptr, adj, _ = to_csr(g["n"], g["edges"])
    value, delta = O.pr(ptr, adj, acc=1e-12)
    np.testing.assert_allclose(value, g["pr"], rtol=1e-10)
    assert (delta <= 1e-12).all()


# ------------------------------------------------------------ networkx pins
nx = pytest.importorskip("networkx")


def nx_graph(ptr, adj, w):
    G = nx.MultiDiGraph()
    n = len(ptr) - 1
    G.add_nodes_from(range(n))
    for u in range(n):
        for e in range(int(ptr[u]), int(ptr[u + 1])):
            G.add_edge(u, int(adj[e]), weight=int(w[e]))
    return G


@pytest.mark.parametrize("seed", range(6))
def test_oracle_vs_networkx(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 300))
    m = int(rng.integers(0, 4 * n))
    ptr, adj, w = random_graph(rng, n, m)
    G = nx_graph(ptr, adj, w)
    src = int(rng.integers(0, n))
    # BFS levels
    lv = nx.single_source_shortest_path_length(G, src)
    exp = np.full(n, INF, np.uint32)
    for v, d in lv.items():
        exp[v] = d
    assert (O.bfs(ptr, adj, src) == exp).all()
    # Dijkstra (networkx uses the minimum parallel edge)
    ds = nx.single_source_dijkstra_path_length(G, src, weight="weight")
    exp = np.full(n, INF, np.uint32)
    for v, d in ds.items():
        exp[v] = d
    assert (O.sssp(ptr, adj, w, src) == exp).all()
    # CC on a directed graph: min ancestor id
    exp = np.array([min([v] + list(nx.ancestors(G, v))) for v in range(n)], np.uint32)
    assert (O.cc(ptr, adj) == exp).all()


@pytest.mark.parametrize("seed", range(4))
def test_oracle_cc_symmetric_equals_components(seed):
    rng = np.random.default_rng(100 + seed)
    n = 200
    m = 150
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    edges = [(int(a), int(b), 1) for a, b in zip(s, d)] + [(int(b), int(a), 1) for a, b in zip(s, d)]
    ptr, adj, w = to_csr(n, edges)
    G = nx.Graph()
    G.add_nodes_from(range(n))
    G.add_edges_from((int(a), int(b)) for a, b in zip(s, d))
    exp = np.zeros(n, np.uint32)
    for comp in nx.connected_components(G):
        exp[list(comp)] = min(comp)
    assert (O.cc(ptr, adj) == exp).all()


# ------------------------------------------------------------ brute force
def all_simple_paths_best(n, edges, src, combine, better, start, unreached):
    """Exhaustive search over simple paths (tiny n only)."""
    best = [unreached] * n
    best[src] = start
    adjl = {}
    for u, v, w in edges:
        adjl.setdefault(u, []).append((v, w))

    def dfs(u, val, seen):
        for v, w in adjl.get(u, []):
            if v in seen:
                continue
            c = combine(val, w)
            if better(c, best[v]):
                best[v] = c
            dfs(v, c, seen | {v})

    dfs(src, start, {src})
    return best


@pytest.mark.parametrize("seed", range(10))
def test_oracle_sssp_sswp_brute_force(seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(2, 8))
    m = int(rng.integers(0, 14))
    edges = [(int(rng.integers(0, n)), int(rng.integers(0, n)), int(rng.integers(1, 65))) for _ in range(m)]
    ptr, adj, w = to_csr(n, edges)
    exp = all_simple_paths_best(n, edges, 0, lambda a, b: a + b, lambda a, b: a < b, 0, INF)
    assert O.sssp(ptr, adj, w, 0).tolist() == exp
    exp = all_simple_paths_best(n, edges, 0, min, lambda a, b: a > b, INF, 0)
    assert O.sswp(ptr, adj, w, 0).tolist() == exp


# ---------------------------------------------------------- PR closed form
@pytest.mark.parametrize("seed", range(4))
def test_oracle_pr_linear_solve(seed):
    rng = np.random.default_rng(7 + seed)
    n = int(rng.integers(20, 120))
    ptr, adj, _ = random_graph(rng, n, 4 * n)
    deg = np.diff(ptr).astype(np.float64)
    M = np.zeros((n, n))
    for u in range(n):
        for e in range(int(ptr[u]), int(ptr[u + 1])):
            M[int(adj[e]), u] += 0.85 / deg[u]
    x = np.linalg.solve(np.eye(n) - M, np.full(n, 0.15))
    acc = 1e-9
    value, delta = O.pr(ptr, adj, acc)
    assert (delta <= acc).all()
    # invariant of delta-push: (I - M) value + delta = 0.15 at every step
    np.testing.assert_allclose((np.eye(n) - M) @ value + delta, 0.15, atol=1e-12)
    # hence x - value = (I - M)^-1 delta, and since M's columns sum to <= 0.85,
    # ||x - value||_1 <= ||delta||_1 / 0.15
    assert np.sum(np.abs(x - value)) <= np.sum(delta) / 0.15 * (1 + 1e-9) + 1e-12
    # a dropped damping factor is caught by the closed form
    bad = np.linalg.solve(np.eye(n) - M / 0.85, np.full(n, 0.15)) if np.linalg.cond(np.eye(n) - M / 0.85) < 1e8 else None
    if bad is not None:
        assert np.max(np.abs(bad - value)) > 1e-3


# ---------------------------------------------------- generator/partitioner
@pytest.mark.parametrize("seed", range(5))
def test_oracle_generate_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 5000))
    act = (rng.random(n) < rng.random()).astype(np.uint8)
    deg = rng.integers(0, 50, n).astype(np.uint32)
    an, ap = O.generate(act, deg)
    idx = np.nonzero(act)[0]
    assert (an == idx).all()
    exp = np.concatenate([[0], np.cumsum(deg[idx].astype(np.uint64))])
    assert (ap == exp).all()


@pytest.mark.parametrize("seed", range(5))
def test_oracle_partition_invariants(seed):
    rng = np.random.default_rng(seed)
    a = int(rng.integers(1, 3000))
    deg = rng.integers(0, 200, a).astype(np.uint64)
    deg[rng.integers(0, a)] = 5000  # a vertex larger than the budget
    ap = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    budget = 1000
    rows = O.partition(ap, budget)
    # contiguous cover in order
    assert rows[0, 0] == 0
    assert (rows[1:, 0] == rows[:-1, 0] + rows[:-1, 1]).all()
    assert rows[-1, 0] + rows[-1, 1] == a
    for fn, cnt, fe, ec in rows:
        assert fe == ap[fn] and ec == ap[fn + cnt] - ap[fn]
        assert ec <= budget or cnt == 1
        # largest fit: adding the next vertex would exceed the budget
        if fn + cnt < a:
            assert ap[fn + cnt + 1] - ap[fn] > budget


def test_planted_mistake_is_caught():
    """A wrong relaxation (max instead of min) in a copy of the SSSP rule
    disagrees with the golden answer, so the golden check has teeth."""
    g = load_hand()["detour"]
    ptr, adj, w = to_csr(g["n"], g["edges"])
    dist = np.full(g["n"], INF, np.int64)
    dist[0] = 0
    for _ in range(g["n"]):
        for u in range(g["n"]):
            if dist[u] == INF:
                continue
            for e in range(int(ptr[u]), int(ptr[u + 1])):
                v = int(adj[e])
                cand = dist[u] + int(w[e])
                if dist[v] == INF or cand > dist[v]:  # planted: keeps the longer path
                    if cand < 10**6:
                        dist[v] = cand
    assert dist.tolist() != g["sssp"]


# ------------------------------------------------------------ OpenMP build
@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("algo", ["bfs", "sssp", "sswp", "cc"])
def test_omp_baseline_matches_oracle(seed, algo):
    rng = np.random.default_rng(50 + seed)
    n = 3000
    ptr, adj, w = random_graph(rng, n, 6 * n)
    got, rounds, edges = O.omp_label(algo, ptr, adj, w, 0)
    assert (got == O.expected(algo, ptr, adj, w, 0)).all()
    assert rounds >= 1


def test_omp_pr_within_tolerance():
    rng = np.random.default_rng(3)
    ptr, adj, _ = random_graph(rng, 2000, 10000)
    v32, _, _ = O.omp_pr(ptr, adj, 1e-7)
    v64, _ = O.pr(ptr, adj, 1e-7)
    assert np.max(np.abs(v32 - v64) / v64) < 1e-4
