// subgraph.hpp -- Subgraph<E>: the compacted active subgraph and the device
// edge budget (drop-in for shared/subgraph.cuh:9-29).
//
// The active list and its offsets live on the device (d_activeNodes,
// d_activeNodesPointer, written by SubgraphGenerator).  The device edge
// budget is split into `numBuffers` partition buffers so the transfer of
// partition i+1 overlaps the processing of partition i; max_partition_size
// is the capacity of one buffer in edges.  SubgraphGenerator::generate()
// leaves the host mirrors valid as the reference does: activeNodes and
// activeNodesPointer (numActiveNodes + 1 offsets) and activeEdgeList, the
// active vertices' edges gathered into pinned memory sized to the active
// edges (grown when a later iteration has more).  The engine's own loop
// keeps everything on the device and never fills them.
#ifndef SUBWAY_SUBGRAPH_HPP
#define SUBWAY_SUBGRAPH_HPP

#include "globals.hpp"

template <class E>
class Subgraph {
public:
    uint num_nodes;
    ull num_edges;
    uint numActiveNodes;
    ull numActiveEdges;

    uint *activeNodes;          // host mirror (pinned), filled by copyToHost()
    ull *activeNodesPointer;    // host mirror (pinned), numActiveNodes+1 entries
    E *activeEdgeList;          // pinned: the active vertices' edges after generate()

    uint *d_activeNodes;
    ull *d_activeNodesPointer;
    E *d_activeEdgeList;        // start of the device edge budget (buffer 0)

    ull max_partition_size;     // edges per device buffer
    ull budget_bytes;           // device bytes of all edge buffers
    int numBuffers;
    E *d_buffers[2];

    // Reference sizing: 0.9 x (device memory - 32 B per vertex)
    // (shared/subgraph.cu:24-28, computed in 64-bit), one buffer.
    Subgraph(uint num_nodes, ull num_edges);
    // Explicit device budget in bytes, split into numBuffers (1 or 2) buffers.
    Subgraph(uint num_nodes, ull num_edges, ull budgetBytes, int numBuffers = 2);
    ~Subgraph();
    Subgraph(const Subgraph &) = delete;
    Subgraph &operator=(const Subgraph &) = delete;

    void copyToHost();              // active list + offsets -> host mirrors
    E *ensureActiveEdgeList(ull edges);   // activeEdgeList with room for `edges` edges
    E *stagingBuffer(int b);        // pinned staging buffer b of the host-gather transfers (on demand)

private:
    void init(ull budgetBytes, int nbuf);
    E *staging[2];
    ull activeEdgeCap = 0;
};

#endif  // SUBWAY_SUBGRAPH_HPP
