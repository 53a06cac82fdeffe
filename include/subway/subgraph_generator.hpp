// subgraph_generator.hpp -- SubgraphGenerator<E> (drop-in for
// shared/subgraph_generator.cuh:13-30).
//
// generate() runs the gfx950 selection (activeness flags label1 || label2,
// or delta > acc, reduced per 16384-vertex chunk, the chunk totals scanned,
// the chunks compacted with wave64 + workgroup scans) into d_activeNodes /
// d_activeNodesPointer (64-bit offsets), and -- as the reference does on
// return (shared/subgraph_generator.cu:143-197) -- leaves numActiveNodes,
// numActiveEdges, the host activeNodes / activeNodesPointer and the gathered
// host activeEdgeList valid.  enqueue() is the device-only half the
// engine's loop uses (no host copies; partitions are gathered when moved).
#ifndef SUBWAY_SUBGRAPH_GENERATOR_HPP
#define SUBWAY_SUBGRAPH_GENERATOR_HPP

#include <hip/hip_runtime.h>

#include "globals.hpp"
#include "graph.hpp"
#include "subgraph.hpp"

namespace subway { struct ScanScratch; }

template <class E>
class SubgraphGenerator {
public:
    hipStream_t stream = nullptr;   // stream the generator kernel runs on

    explicit SubgraphGenerator(Graph<E> &graph);
    explicit SubgraphGenerator(GraphPR<E> &graph);
    ~SubgraphGenerator();
    SubgraphGenerator(const SubgraphGenerator &) = delete;
    SubgraphGenerator &operator=(const SubgraphGenerator &) = delete;

    void generate(Graph<E> &graph, Subgraph<E> &subgraph);
    void generate(GraphPR<E> &graph, Subgraph<E> &subgraph, float acc);

    // Enqueue only (no host synchronisation): totals stay on the device in
    // the control block returned by deviceTotals(): u64 words [count, edge
    // total, error]; error (low 32 bits) is nonzero when the kernel's bounded
    // lookback wait gave up, and then count must be read as zero.
    void enqueue(Graph<E> &graph, Subgraph<E> &subgraph);
    void enqueue(GraphPR<E> &graph, Subgraph<E> &subgraph, float acc);
    const unsigned long long *deviceTotals() const;   // [count, total, error]
    // Scan scratch of the generator kernel (other selection rules, e.g. the
    // batched per-source masks, run through it so deviceTotals() holds them).
    subway::ScanScratch &scanScratch() { return *scratch; }

private:
    subway::ScanScratch *scratch;
    ull *h_totals;   // pinned readback
    void readback(Subgraph<E> &subgraph);
};

#endif  // SUBWAY_SUBGRAPH_GENERATOR_HPP
