// subgraph_generator.hpp -- SubgraphGenerator<E> (drop-in for
// shared/subgraph_generator.cuh:13-30).
//
// generate() runs one single-pass gfx950 kernel: activeness flags
// (label1 || label2, or delta > acc), a wave64 + workgroup scan of the flags
// and of the masked degrees, a decoupled-lookback prefix across workgroups,
// and the compaction into d_activeNodes / d_activeNodesPointer (64-bit
// offsets).  Like the reference it returns with numActiveNodes set; unlike
// the reference it moves no edge data -- the edges of each partition are
// gathered when the partition is transferred.
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
