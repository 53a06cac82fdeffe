// partitioner.cpp -- host side of Partitioner<E>
// (reference: shared/partitioner.cu:4-104).
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "subway/gpu_error_check.hpp"
#include "subway/partitioner.hpp"

using namespace subway;

static constexpr uint32_t MAX_ROWS = 4096;
static constexpr uint32_t FIRST_ROWS = 64;

template <class E>
Partitioner<E>::Partitioner() : numPartitions(0), numActiveNodes(0), numActiveEdges(0), d_rows(nullptr), d_out(nullptr), h_rows(nullptr), h_out(nullptr)
{
    SUBWAY_CHECK(hipMalloc(&d_rows, MAX_ROWS * sizeof(PartRow)));
    SUBWAY_CHECK(hipMalloc(&d_out, 4 * sizeof(ull)));
    SUBWAY_CHECK(hipHostMalloc(&h_rows, MAX_ROWS * sizeof(PartRow), 0));
    SUBWAY_CHECK(hipHostMalloc((void **)&h_out, 4 * sizeof(ull), 0));
    reset();
}

template <class E>
Partitioner<E>::~Partitioner()
{
    if (d_rows) (void)hipFree(d_rows);
    if (d_out) (void)hipFree(d_out);
    if (h_rows) (void)hipHostFree(h_rows);
    if (h_out) (void)hipHostFree(h_out);
}

template <class E>
void Partitioner<E>::reset()
{
    fromNode.clear();
    fromEdge.clear();
    partitionNodeSize.clear();
    partitionEdgeSize.clear();
    firstVertex.clear();
    lastVertex.clear();
    numPartitions = 0;
}

template <class E>
void Partitioner<E>::partition(Subgraph<E> &subgraph, uint numActiveNodes)
{
    run(subgraph, nullptr, numActiveNodes, subgraph.max_partition_size);
}

template <class E>
void Partitioner<E>::partition(Subgraph<E> &subgraph, const unsigned long long *d_numActive,
                               unsigned long long budgetEdges)
{
    run(subgraph, d_numActive, 0, budgetEdges);
}

template <class E>
void Partitioner<E>::enqueueRows(Subgraph<E> &subgraph, const unsigned long long *d_numActive, ull hostCount,
                                 ull budgetEdges, ull start)
{
    launch_partition(subgraph.d_activeNodesPointer, subgraph.d_activeNodes, (const ull *)d_numActive, hostCount,
                     budgetEdges, start, MAX_ROWS, (PartRow *)d_rows, (ull *)d_out, stream);
    // the first rows travel with the count; a long table takes a second copy
    SUBWAY_CHECK(hipMemcpyAsync(h_out, d_out, 4 * sizeof(ull), hipMemcpyDeviceToHost, stream));
    SUBWAY_CHECK(hipMemcpyAsync(h_rows, d_rows, FIRST_ROWS * sizeof(PartRow), hipMemcpyDeviceToHost, stream));
}

// Reads the rows of the last enqueueRows; returns true when the table
// continues past MAX_ROWS (h_out[1] is then the next start).
template <class E>
bool Partitioner<E>::collectRows()
{
    SUBWAY_CHECK(hipStreamSynchronize(stream));
    const PartRow *rows = (const PartRow *)h_rows;
    ull got = h_out[0];
    numActiveNodes = h_out[2];
    numActiveEdges = h_out[3];
    if (got > FIRST_ROWS) {
        SUBWAY_CHECK(hipMemcpyAsync((PartRow *)h_rows + FIRST_ROWS, (PartRow *)d_rows + FIRST_ROWS,
                                    (got - FIRST_ROWS) * sizeof(PartRow), hipMemcpyDeviceToHost, stream));
        SUBWAY_CHECK(hipStreamSynchronize(stream));
    }
    for (ull r = 0; r < got; r++) {
        fromNode.push_back(rows[r].fromNode);
        partitionNodeSize.push_back(rows[r].nodeCount);
        fromEdge.push_back(rows[r].fromEdge);
        partitionEdgeSize.push_back(rows[r].edgeCount);
        firstVertex.push_back((uint)rows[r].firstId);
        lastVertex.push_back((uint)rows[r].lastId);
    }
    return got >= MAX_ROWS;
}

template <class E>
void Partitioner<E>::finish(Subgraph<E> &subgraph, const unsigned long long *d_numActive, ull hostCount,
                            ull budgetEdges)
{
    while (collectRows()) enqueueRows(subgraph, d_numActive, hostCount, budgetEdges, h_out[1]);
    numPartitions = (uint)fromNode.size();
}

template <class E>
void Partitioner<E>::run(Subgraph<E> &subgraph, const unsigned long long *d_numActive, unsigned long long hostCount,
                         unsigned long long budgetEdges)
{
    reset();
    numActiveNodes = hostCount;
    numActiveEdges = 0;
    if (!d_numActive && hostCount == 0) return;
    enqueueRows(subgraph, d_numActive, hostCount, budgetEdges, 0);
    finish(subgraph, d_numActive, hostCount, budgetEdges);
}

template <class E>
void Partitioner<E>::enqueue(Subgraph<E> &subgraph, const unsigned long long *d_numActive,
                             unsigned long long budgetEdges)
{
    enqueueRows(subgraph, d_numActive, 0, budgetEdges, 0);
}

template <class E>
void Partitioner<E>::collect(Subgraph<E> &subgraph, const unsigned long long *d_numActive,
                             unsigned long long budgetEdges)
{
    reset();
    numActiveNodes = 0;
    numActiveEdges = 0;
    finish(subgraph, d_numActive, 0, budgetEdges);
}

// shared/partitioner.cu:103-104
template class Partitioner<OutEdge>;
template class Partitioner<OutEdgeWeighted>;
