// subgraph_generator.cpp -- host side of SubgraphGenerator<E>
// (reference: shared/subgraph_generator.cu:100-285).
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>
#include <string>

#include "host_util.hpp"

#include "kernels.hpp"
#include "subway/gpu_error_check.hpp"
#include "subway/subgraph_generator.hpp"

using namespace subway;

template <class E>
SubgraphGenerator<E>::SubgraphGenerator(Graph<E> &graph) : scratch(new ScanScratch), h_totals(nullptr)
{
    alloc_scan_scratch(*scratch, ((ull)graph.num_nodes + GEN_TILE - 1) / GEN_TILE);
    SUBWAY_CHECK(hipHostMalloc((void **)&h_totals, 3 * sizeof(ull), 0));
}

template <class E>
SubgraphGenerator<E>::SubgraphGenerator(GraphPR<E> &graph) : scratch(new ScanScratch), h_totals(nullptr)
{
    alloc_scan_scratch(*scratch, ((ull)graph.num_nodes + GEN_TILE - 1) / GEN_TILE);
    SUBWAY_CHECK(hipHostMalloc((void **)&h_totals, 3 * sizeof(ull), 0));
}

template <class E>
SubgraphGenerator<E>::~SubgraphGenerator()
{
    free_scan_scratch(*scratch);
    delete scratch;
    if (h_totals) (void)hipHostFree(h_totals);
}

template <class E>
const unsigned long long *SubgraphGenerator<E>::deviceTotals() const
{
    return (const unsigned long long *)&scratch->ctl->count;
}

template <class E>
void SubgraphGenerator<E>::enqueue(Graph<E> &graph, Subgraph<E> &subgraph)
{
    launch_generate_labels(graph.num_nodes, (const uint8_t *)graph.d_label1, (const uint8_t *)graph.d_label2,
                           graph.d_outDegree, subgraph.d_activeNodes, subgraph.d_activeNodesPointer,
                           *scratch, stream);
}

template <class E>
void SubgraphGenerator<E>::enqueue(GraphPR<E> &graph, Subgraph<E> &subgraph, float acc)
{
    launch_generate_pr(graph.num_nodes, graph.d_delta, acc, graph.d_outDegree, subgraph.d_activeNodes,
                       subgraph.d_activeNodesPointer, *scratch, stream);
}

template <class E>
void SubgraphGenerator<E>::readback(Subgraph<E> &subgraph)
{
    // [count, total, error] are adjacent u64 words of the control block
    SUBWAY_CHECK(hipMemcpyAsync(h_totals, &scratch->ctl->count, 3 * sizeof(ull), hipMemcpyDeviceToHost, stream));
    SUBWAY_CHECK(hipStreamSynchronize(stream));
    const uint32_t err = (uint32_t)h_totals[2];
    if (err)
        throw subway::Error(subway::ERR_INTERNAL,
                            std::string("generator: device wait gave up (") + subway::wait_error_name(err) + ")");
    subgraph.numActiveNodes = (uint)h_totals[0];
    subgraph.numActiveEdges = h_totals[1];
}

// The host side of the reference's generate (shared/subgraph_generator.cu:
// 143-197): the active list and its offsets copied to the host mirrors, and
// every active vertex's edges gathered from the host CSR into the pinned
// activeEdgeList at its offset, by host threads over blocks of the list
// (the whole list: no dropped tail).
template <class E>
static void gather_active_edges(const subway::HostCsr<E> &csr, Subgraph<E> &sg)
{
    sg.copyToHost();
    E *out = sg.ensureActiveEdgeList(sg.numActiveEdges);
    const uint *ids = sg.activeNodes;
    const ull *at = sg.activeNodesPointer;
    subway::parallel_for(sg.numActiveNodes, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t i = b; i < e; i++) {
            const ull p0 = csr.ptr[ids[i]], d = csr.ptr[ids[i] + 1] - p0;
            if (csr.edges) {
                std::memcpy(out + at[i], csr.edges + p0, d * sizeof(E));
            } else {
                for (ull k = 0; k < d; k++) out[at[i] + k] = csr.edgeAt(p0 + k);
            }
        }
    });
}

template <class E>
void SubgraphGenerator<E>::generate(Graph<E> &graph, Subgraph<E> &subgraph)
{
    enqueue(graph, subgraph);
    readback(subgraph);
    gather_active_edges(*graph.csr, subgraph);
}

template <class E>
void SubgraphGenerator<E>::generate(GraphPR<E> &graph, Subgraph<E> &subgraph, float acc)
{
    enqueue(graph, subgraph, acc);
    readback(subgraph);
    gather_active_edges(*graph.csr, subgraph);
}

// shared/subgraph_generator.cu:287-288
template class SubgraphGenerator<OutEdge>;
template class SubgraphGenerator<OutEdgeWeighted>;
