// partitioner.hpp -- Partitioner<E> (drop-in for shared/partitioner.cuh:9-22).
//
// partition() cuts the active list into contiguous ranges whose edges fit
// one device buffer (subgraph.max_partition_size).  The split points are
// found on the device by a one-wave 64-ary search over the offsets (exact
// largest fit), and read back as one small table.  A vertex whose degree
// exceeds the buffer forms a partition of its own; the engine streams such
// a vertex in buffer-sized chunks (the reference loops forever there,
// shared/partitioner.cu:47-70).
#ifndef SUBWAY_PARTITIONER_HPP
#define SUBWAY_PARTITIONER_HPP

#include <hip/hip_runtime.h>

#include <vector>

#include "globals.hpp"
#include "subgraph.hpp"

template <class E>
class Partitioner {
public:
    uint numPartitions;
    std::vector<ull> fromNode;
    std::vector<ull> fromEdge;
    std::vector<ull> partitionNodeSize;
    std::vector<ull> partitionEdgeSize;
    std::vector<uint> firstVertex;   // id of the first active vertex of each partition
    std::vector<uint> lastVertex;    // id of the last active vertex of each partition
    ull numActiveNodes;              // active count seen by the last partition() call
    ull numActiveEdges;              // active edge total seen by the last partition() call
    hipStream_t stream = nullptr;

    Partitioner();
    ~Partitioner();
    Partitioner(const Partitioner &) = delete;
    Partitioner &operator=(const Partitioner &) = delete;

    void partition(Subgraph<E> &subgraph, uint numActiveNodes);
    // Variant reading the active count from device memory (no extra sync).
    void partition(Subgraph<E> &subgraph, const unsigned long long *d_numActive,
                   unsigned long long budgetEdges);
    // The same in two halves, so the first can be recorded in a HIP graph:
    // enqueue() issues the partition kernel for the first rows and the copies
    // of the table to the host on `stream` (no synchronisation); collect()
    // synchronises the stream, reads the table and finishes a table longer
    // than one launch holds.
    void enqueue(Subgraph<E> &subgraph, const unsigned long long *d_numActive, unsigned long long budgetEdges);
    void collect(Subgraph<E> &subgraph, const unsigned long long *d_numActive, unsigned long long budgetEdges);
    void reset();

private:
    void *d_rows;
    unsigned long long *d_out;
    void *h_rows;
    unsigned long long *h_out;
    void run(Subgraph<E> &subgraph, const unsigned long long *d_numActive, unsigned long long hostCount,
             unsigned long long budgetEdges);
    void enqueueRows(Subgraph<E> &subgraph, const unsigned long long *d_numActive, unsigned long long hostCount,
                     unsigned long long budgetEdges, unsigned long long start);
    bool collectRows();
    void finish(Subgraph<E> &subgraph, const unsigned long long *d_numActive, unsigned long long hostCount,
                unsigned long long budgetEdges);
};

#endif  // SUBWAY_PARTITIONER_HPP
