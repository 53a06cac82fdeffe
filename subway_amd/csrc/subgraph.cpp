// subgraph.cpp -- Subgraph<E> buffers and the device edge budget
// (reference: shared/subgraph.cu:9-42).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "subway/gpu_error_check.hpp"
#include "subway/subgraph.hpp"

using subway::Error;

template <class E>
Subgraph<E>::Subgraph(uint n, ull m)
    : num_nodes(n), num_edges(m), numActiveNodes(0), numActiveEdges(0), activeNodes(nullptr),
      activeNodesPointer(nullptr), activeEdgeList(nullptr), d_activeNodes(nullptr),
      d_activeNodesPointer(nullptr), d_activeEdgeList(nullptr), max_partition_size(0), budget_bytes(0),
      numBuffers(1), d_buffers{nullptr, nullptr}, staging{nullptr, nullptr}
{
    int dev = 0;
    SUBWAY_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SUBWAY_CHECK(hipGetDeviceProperties(&prop, dev));
    // shared/subgraph.cu:24: 0.9 * (totalGlobalMem - 8*4*num_nodes), in 64-bit
    ull reserve = 32ull * (ull)n;
    ull total = (ull)prop.totalGlobalMem;
    ull budget = total > reserve ? (ull)(0.9 * (double)(total - reserve)) : 0;
    init(budget, 1);
}

template <class E>
Subgraph<E>::Subgraph(uint n, ull m, ull budgetBytes, int nbuf)
    : num_nodes(n), num_edges(m), numActiveNodes(0), numActiveEdges(0), activeNodes(nullptr),
      activeNodesPointer(nullptr), activeEdgeList(nullptr), d_activeNodes(nullptr),
      d_activeNodesPointer(nullptr), d_activeEdgeList(nullptr), max_partition_size(0), budget_bytes(0),
      numBuffers(1), d_buffers{nullptr, nullptr}, staging{nullptr, nullptr}
{
    init(budgetBytes, nbuf);
}

template <class E>
void Subgraph<E>::init(ull budgetBytes, int nbuf)
{
    if (nbuf != 1 && nbuf != 2) throw Error(subway::ERR_ARG, "numBuffers must be 1 or 2");
    numBuffers = nbuf;
    // One buffer holds at most 2^31 - 1 edges (work-space offsets are
    // 32-bit) and a whole number of advance chunks, so oversized vertices can
    // be streamed through it chunk by chunk.
    ull per = budgetBytes / (ull)nbuf / sizeof(E);
    per = std::min<ull>(per, 0x7FFFFFFFull);
    // never allocate more than the whole graph needs
    ull need = std::max<ull>(num_edges, (ull)subway::ADV_CHUNK);
    need = (need + subway::ADV_CHUNK - 1) / subway::ADV_CHUNK * subway::ADV_CHUNK;
    per = std::min(per, need);
    per -= per % subway::ADV_CHUNK;
    if (per < (ull)subway::ADV_CHUNK)
        throw Error(subway::ERR_BUDGET, "device edge budget of " + std::to_string(budgetBytes) +
                                            " bytes is below the minimum of " +
                                            std::to_string((ull)nbuf * subway::ADV_CHUNK * sizeof(E)));
    max_partition_size = per;
    budget_bytes = per * sizeof(E) * (ull)nbuf;
    SUBWAY_CHECK(hipMalloc(&d_activeNodes, ((size_t)num_nodes + 1) * sizeof(uint)));
    SUBWAY_CHECK(hipMalloc(&d_activeNodesPointer, ((size_t)num_nodes + 1) * sizeof(ull)));
    for (int b = 0; b < nbuf; b++) {
        hipError_t e = hipMalloc(&d_buffers[b], (size_t)per * sizeof(E));
        if (e != hipSuccess)
            throw Error(subway::ERR_BUDGET, "device edge buffer of " + std::to_string(per * sizeof(E)) +
                                                " bytes could not be allocated: " + hipGetErrorString(e));
    }
    d_activeEdgeList = d_buffers[0];
}

template <class E>
Subgraph<E>::~Subgraph()
{
    if (activeNodes) (void)hipHostFree(activeNodes);
    if (activeNodesPointer) (void)hipHostFree(activeNodesPointer);
    if (activeEdgeList) (void)hipHostFree(activeEdgeList);
    for (int b = 0; b < 2; b++) {
        if (staging[b]) (void)hipHostFree(staging[b]);
        if (d_buffers[b]) (void)hipFree(d_buffers[b]);
    }
    if (d_activeNodes) (void)hipFree(d_activeNodes);
    if (d_activeNodesPointer) (void)hipFree(d_activeNodesPointer);
}

template <class E>
void Subgraph<E>::copyToHost()
{
    if (!activeNodes) {
        SUBWAY_CHECK(hipHostMalloc((void **)&activeNodes, ((size_t)num_nodes + 1) * sizeof(uint), 0));
        SUBWAY_CHECK(hipHostMalloc((void **)&activeNodesPointer, ((size_t)num_nodes + 1) * sizeof(ull), 0));
    }
    SUBWAY_CHECK(hipMemcpy(activeNodes, d_activeNodes, (size_t)numActiveNodes * sizeof(uint),
                           hipMemcpyDeviceToHost));
    SUBWAY_CHECK(hipMemcpy(activeNodesPointer, d_activeNodesPointer,
                           ((size_t)numActiveNodes + 1) * sizeof(ull), hipMemcpyDeviceToHost));
}

template <class E>
E *Subgraph<E>::stagingBuffer(int b)
{
    if (b < 0 || b >= numBuffers) throw Error(subway::ERR_ARG, "staging buffer index out of range");
    if (!staging[b])
        SUBWAY_CHECK(hipHostMalloc((void **)&staging[b], (size_t)max_partition_size * sizeof(E), 0));
    return staging[b];
}

template <class E>
E *Subgraph<E>::ensureActiveEdgeList(ull edges)
{
    if (edges > activeEdgeCap || !activeEdgeList) {
        if (activeEdgeList) SUBWAY_CHECK(hipHostFree(activeEdgeList));
        activeEdgeList = nullptr;
        // grow in steps of an eighth of the graph, never past it
        ull cap = std::max<ull>(edges, std::min<ull>(num_edges, activeEdgeCap + num_edges / 8 + 1));
        cap = std::max<ull>(cap, 1);
        SUBWAY_CHECK(hipHostMalloc((void **)&activeEdgeList, (size_t)cap * sizeof(E), 0));
        activeEdgeCap = cap;
    }
    return activeEdgeList;
}

// shared/subgraph.cu:44-45
template class Subgraph<OutEdge>;
template class Subgraph<OutEdgeWeighted>;
