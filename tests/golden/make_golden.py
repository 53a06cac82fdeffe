"""Writes the golden fixtures of tests/golden/.

Inputs:
  readme_graph.el / readme_graph.wel -- the two sample graphs of the
      reference's README (README.md:16-30), the only sample inputs it ships.
  hand graphs -- small graphs chosen so that a dropped term, a wrong
      comparison direction or a wrong initial value changes the answer.
Expected outputs are derived by hand (derivations inline) and stored as
JSON; the expected .bcsr/.bwcsr bytes are written from the format layout
[u32 n][u32 m][u32 nodePointer x n][edges] (shared/graph.cu:39-51,
tools/converter.cpp:88-93,157-162) with the stable by-source edge order of
tools/converter.cpp:64-83.

Run: python tests/golden/make_golden.py
"""
import json
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))
INF = 4294967294

README_EL = [(0, 1), (0, 3), (2, 3), (1, 2)]
README_WEL = [(0, 1, 26), (0, 3, 33), (2, 3, 40), (1, 2, 10)]

# PageRank closed form for the README graph (no cycles): x = 0.15 + 0.85 * sum(x_u / deg_u)
# deg(0)=2, deg(1)=1, deg(2)=1, deg(3)=0
_x0 = 0.15
_x1 = 0.15 + 0.85 * _x0 / 2
_x2 = 0.15 + 0.85 * _x1
_x3 = 0.15 + 0.85 * (_x0 / 2 + _x2)

GRAPHS = {
    "readme": {
        "n": 4,
        "edges": README_WEL,
        "source": 0,
        # BFS: 0 ->1, 0->3 (level 1), 1->2 (level 2)
        "bfs": [0, 1, 2, 1],
        # SSSP: 0->1 = 26, 0->1->2 = 36, 0->3 = 33 (beats 0->1->2->3 = 76)
        "sssp": [0, 26, 36, 33],
        # SSWP: source INF; 1 via 26; 2 via min(26,10)=10; 3: max(33, min(10,40)) = 33
        "sswp": [INF, 26, 10, 33],
        # CC: every vertex is reachable from 0
        "cc": [0, 0, 0, 0],
        "pr": [_x0, _x1, _x2, _x3],
    },
    # Two-hop path cheaper than the direct edge; an unreachable vertex (4);
    # a vertex (5) reaching into the graph but not reachable from it.
    "detour": {
        "n": 6,
        "edges": [(0, 1, 10), (0, 2, 1), (2, 1, 2), (1, 3, 1), (2, 3, 7), (5, 3, 1), (5, 0, 50)],
        "source": 0,
        "bfs": [0, 1, 1, 2, INF, INF],
        # 1: min(10, 1+2) = 3; 3: min(3+1, 1+7) = 4
        "sssp": [0, 3, 1, 4, INF, INF],
        # 1: max(10, min(1,2)=1) = 10; 2: 1; 3: max(min(10,1)=1, min(1,7)=1) = 1
        "sswp": [INF, 10, 1, 1, 0, 0],
        # min ancestor: 5 reaches 0,1,2,3 but 0 < 5; 4 isolated; 5 only itself
        "cc": [0, 0, 0, 0, 4, 5],
    },
    # Directed cycle 3->4->5->3 entered from 1; cycle 6<->7 with no entry.
    "cycles": {
        "n": 8,
        "edges": [(1, 3, 5), (3, 4, 5), (4, 5, 5), (5, 3, 5), (6, 7, 2), (7, 6, 2), (0, 1, 64), (2, 2, 1)],
        "source": 1,
        "bfs": [INF, 0, INF, 1, 2, 3, INF, INF],
        "sssp": [INF, 0, INF, 5, 10, 15, INF, INF],
        "sswp": [0, INF, 0, 5, 5, 5, 0, 0],
        # 0 reaches 1,3,4,5; 2 has only a self-loop; 6 and 7 reach each other
        "cc": [0, 0, 2, 0, 0, 0, 6, 6],
    },
    # Widest path differs from shortest path: 0->1->3 is wide (50, 40),
    # 0->2->3 is short (1, 1).
    "wide": {
        "n": 4,
        "edges": [(0, 1, 50), (1, 3, 40), (0, 2, 1), (2, 3, 1)],
        "source": 0,
        "bfs": [0, 1, 1, 2],
        "sssp": [0, 50, 1, 2],
        "sswp": [INF, 50, 1, 40],
        "cc": [0, 0, 0, 0],
    },
}


def csr(n, edges):
    by = [[] for _ in range(n)]
    for e in edges:
        by[e[0]].append(e)
    ptr, out = [0], []
    for v in range(n):
        out.extend(by[v])
        ptr.append(len(out))
    return ptr, out


def bcsr_bytes(n, edges, weighted):
    ptr, out = csr(n, edges)
    b = struct.pack("<II", n, len(out)) + struct.pack(f"<{n}I", *ptr[:n])
    for e in out:
        b += struct.pack("<II", e[1], e[2]) if weighted else struct.pack("<I", e[1])
    return b


def main():
    with open(os.path.join(HERE, "readme_graph.el"), "w") as f:
        f.write("".join(f"{s} {d}\n" for s, d in README_EL))
    with open(os.path.join(HERE, "readme_graph.wel"), "w") as f:
        f.write("".join(f"{s} {d} {w}\n" for s, d, w in README_WEL))
    with open(os.path.join(HERE, "readme_graph.expected.bcsr"), "wb") as f:
        f.write(bcsr_bytes(4, [(s, d, 0) for s, d in README_EL], False))
    with open(os.path.join(HERE, "readme_graph.expected.bwcsr"), "wb") as f:
        f.write(bcsr_bytes(4, README_WEL, True))
    with open(os.path.join(HERE, "hand_graphs.json"), "w") as f:
        json.dump(GRAPHS, f, indent=1)


if __name__ == "__main__":
    main()
