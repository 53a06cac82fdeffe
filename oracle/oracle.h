/*
 * oracle.h -- CPU restatement of Subway's hot path, used ONLY as test
 * infrastructure (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg).  Nothing in the product library links or calls this code.
 *
 * Graphs are CSR with 64-bit offsets: ptr[n+1], adj[m] (destination ids),
 * w[m] (uint32 weights, only for weighted algorithms).
 *
 * Parity pins (what checks each function against something independent):
 *   oracle_bfs   - hand graphs with known levels; networkx BFS; invariants
 *   oracle_sssp  - hand graphs; networkx Dijkstra; Bellman-Ford brute force
 *   oracle_sswp  - hand graphs; brute force over all simple paths (n<=8);
 *                  bottleneck fixpoint invariant
 *   oracle_cc    - hand graphs; networkx reachability (min ancestor);
 *                  union-find on symmetric graphs
 *   oracle_pr    - closed form: dense linear solve of x = 0.15 + 0.85 P^T x,
 *                  error bound from the residual threshold
 *   oracle_generate / oracle_partition - brute force in the tests
 */
#ifndef SUBWAY_ORACLE_H
#define SUBWAY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* shared/globals.hpp:27  DIST_INFINITY = UINT_MAX - 1 */
#define ORACLE_INF 4294967294u

void oracle_bfs(uint64_t n, const uint64_t *ptr, const uint32_t *adj,
                uint32_t src, uint32_t *level);
void oracle_sssp(uint64_t n, const uint64_t *ptr, const uint32_t *adj,
                 const uint32_t *w, uint32_t src, uint32_t *dist);
void oracle_sswp(uint64_t n, const uint64_t *ptr, const uint32_t *adj,
                 const uint32_t *w, uint32_t src, uint32_t *width);
void oracle_cc(uint64_t n, const uint64_t *ptr, const uint32_t *adj,
               uint32_t *label);
/* returns the number of vertex pushes performed */
uint64_t oracle_pr(uint64_t n, const uint64_t *ptr, const uint32_t *adj,
                   double init, double damping, double acc, double *value,
                   double *delta);

/* Subgraph generation: flags -> exclusive scans -> compaction.
 * active[v] != 0 marks v; ptr_out has numActive+1 entries. */
uint64_t oracle_generate(uint64_t n, const uint8_t *active,
                         const uint32_t *degree, uint32_t *activeNodes,
                         uint64_t *activeNodesPointer);

/* Greedy contiguous partitioning of the active list with a budget in edges.
 * A vertex whose degree exceeds the budget forms its own partition.
 * Writes up to max_parts (fromNode, nodeCount, fromEdge, edgeCount) rows and
 * returns the number of partitions. */
uint64_t oracle_partition(uint64_t numActive, const uint64_t *activeNodesPointer,
                          uint64_t budget_edges, uint64_t max_parts,
                          uint64_t *fromNode, uint64_t *nodeCount,
                          uint64_t *fromEdge, uint64_t *edgeCount);

#ifdef __cplusplus
}
#endif

#endif
