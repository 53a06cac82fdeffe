// A driver written against include/subway in the order the reference's
// applications use its classes and kernels (subway/sssp-async.cu:38-122,
// subway/bfs-sync.cu:28-89, subway/pr-async.cu:34-112): host-side initial
// values copied to the device, one generate, then per global iteration
// partition -> per partition copy the slice of activeEdgeList -> label
// kernel(s) -> generate.  It uses only the reference-shaped API: the public
// fields, Subgraph / SubgraphGenerator / Partitioner, the host launchers of
// gpu_kernels.hpp and gpuErrorcheck.
//
//   ref_loop <bfs|cc|sssp|sswp|pr>-<sync|async> <graph file> <source> <budget bytes> <output file>
//
// The output file has the reference's "<vertex> <value>" lines.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <iostream>
#include <string>

#include "subway/globals.hpp"
#include "subway/gpu_error_check.hpp"
#include "subway/gpu_kernels.hpp"
#include "subway/graph.hpp"
#include "subway/partitioner.hpp"
#include "subway/subgraph.hpp"
#include "subway/subgraph_generator.hpp"
#include "subway/subway_utilities.hpp"

// a partition's edges must fit the device buffer they are copied into (a
// vertex larger than the buffer has no place in this loop, as in the
// reference)
static void fits(ull edges, ull capacity)
{
    if (edges > capacity) {
        std::cerr << "partition of " << edges << " edges exceeds the device buffer of " << capacity << "\n";
        std::exit(3);
    }
}

template <class E, class G>
static void to_device(G &graph)
{
    gpuErrorcheck(hipMemcpy(graph.d_outDegree, graph.outDegree, graph.num_nodes * sizeof(uint), hipMemcpyHostToDevice));
    gpuErrorcheck(hipMemcpy(graph.d_value, graph.value, graph.num_nodes * sizeof(graph.value[0]),
                            hipMemcpyHostToDevice));
}

// async label algorithms: every vertex labelled for the first generate, then
// only the source; rounds on each partition until no destination improves
template <class E, class Kernel>
static int label_async(const std::string &path, uint src, ull budget, const std::string &out, bool weighted,
                       uint init, uint srcValue, bool allLabelled, Kernel kernel)
{
    Graph<E> graph(path, weighted);
    graph.ReadGraph();
    for (uint v = 0; v < graph.num_nodes; v++) {
        graph.value[v] = allLabelled ? v : init;
        graph.label1[v] = true;
        graph.label2[v] = false;
    }
    if (!allLabelled) graph.value[src] = srcValue;
    to_device<E>(graph);
    gpuErrorcheck(hipMemcpy(graph.d_label1, graph.label1, graph.num_nodes, hipMemcpyHostToDevice));
    gpuErrorcheck(hipMemcpy(graph.d_label2, graph.label2, graph.num_nodes, hipMemcpyHostToDevice));
    Subgraph<E> subgraph(graph.num_nodes, graph.num_edges, budget, 1);
    SubgraphGenerator<E> subgen(graph);
    subgen.generate(graph, subgraph);
    if (!allLabelled) {
        for (uint v = 0; v < graph.num_nodes; v++) graph.label1[v] = (v == src);
        gpuErrorcheck(hipMemcpy(graph.d_label1, graph.label1, graph.num_nodes, hipMemcpyHostToDevice));
    }
    Partitioner<E> partitioner;
    bool finished = false, *d_finished = nullptr;
    gpuErrorcheck(hipMalloc(&d_finished, sizeof(bool)));
    uint iterations = 0;
    while (subgraph.numActiveNodes > 0) {
        iterations++;
        partitioner.partition(subgraph, subgraph.numActiveNodes);
        for (uint i = 0; i < partitioner.numPartitions; i++) {
            fits(partitioner.partitionEdgeSize[i], subgraph.max_partition_size);
            gpuErrorcheck(hipMemcpy(subgraph.d_activeEdgeList, subgraph.activeEdgeList + partitioner.fromEdge[i],
                                    partitioner.partitionEdgeSize[i] * sizeof(E), hipMemcpyHostToDevice));
            mixLabels(subgraph.d_activeNodes, graph.d_label1, graph.d_label2, (uint)partitioner.partitionNodeSize[i],
                      (uint)partitioner.fromNode[i]);
            for (uint round = 1;; round++) {
                finished = true;
                gpuErrorcheck(hipMemcpy(d_finished, &finished, sizeof(bool), hipMemcpyHostToDevice));
                kernel((uint)partitioner.partitionNodeSize[i], (uint)partitioner.fromNode[i],
                       partitioner.fromEdge[i], subgraph.d_activeNodes, subgraph.d_activeNodesPointer,
                       subgraph.d_activeEdgeList, graph.d_outDegree, graph.d_value, d_finished,
                       round % 2 ? graph.d_label1 : graph.d_label2, round % 2 ? graph.d_label2 : graph.d_label1);
                gpuErrorcheck(hipDeviceSynchronize());
                gpuErrorcheck(hipMemcpy(&finished, d_finished, sizeof(bool), hipMemcpyDeviceToHost));
                if (finished) break;
            }
        }
        subgen.generate(graph, subgraph);
    }
    std::cout << "Global iterations = " << iterations << std::endl;
    gpuErrorcheck(hipMemcpy(graph.value, graph.d_value, graph.num_nodes * sizeof(uint), hipMemcpyDeviceToHost));
    utilities::SaveResults(out, graph.value, graph.num_nodes);
    gpuErrorcheck(hipFree(d_finished));
    return 0;
}

// sync label algorithms: label2 of the source (every vertex for CC), then
// per partition moveUpLabels + one kernel launch
template <class E, class Kernel>
static int label_sync(const std::string &path, uint src, ull budget, const std::string &out, bool weighted,
                      uint init, uint srcValue, bool allLabelled, Kernel kernel)
{
    Graph<E> graph(path, weighted);
    graph.ReadGraph();
    for (uint v = 0; v < graph.num_nodes; v++) {
        graph.value[v] = allLabelled ? v : init;
        graph.label1[v] = false;
        graph.label2[v] = allLabelled || v == src;
    }
    if (!allLabelled) graph.value[src] = srcValue;
    to_device<E>(graph);
    gpuErrorcheck(hipMemcpy(graph.d_label1, graph.label1, graph.num_nodes, hipMemcpyHostToDevice));
    gpuErrorcheck(hipMemcpy(graph.d_label2, graph.label2, graph.num_nodes, hipMemcpyHostToDevice));
    Subgraph<E> subgraph(graph.num_nodes, graph.num_edges, budget, 1);
    SubgraphGenerator<E> subgen(graph);
    subgen.generate(graph, subgraph);
    Partitioner<E> partitioner;
    uint iterations = 0;
    while (subgraph.numActiveNodes > 0) {
        iterations++;
        partitioner.partition(subgraph, subgraph.numActiveNodes);
        for (uint i = 0; i < partitioner.numPartitions; i++) {
            fits(partitioner.partitionEdgeSize[i], subgraph.max_partition_size);
            gpuErrorcheck(hipMemcpy(subgraph.d_activeEdgeList, subgraph.activeEdgeList + partitioner.fromEdge[i],
                                    partitioner.partitionEdgeSize[i] * sizeof(E), hipMemcpyHostToDevice));
            moveUpLabels(subgraph.d_activeNodes, graph.d_label1, graph.d_label2, (uint)partitioner.partitionNodeSize[i],
                         (uint)partitioner.fromNode[i]);
            kernel((uint)partitioner.partitionNodeSize[i], (uint)partitioner.fromNode[i], partitioner.fromEdge[i],
                   subgraph.d_activeNodes, subgraph.d_activeNodesPointer, subgraph.d_activeEdgeList, graph.d_outDegree,
                   graph.d_value, graph.d_label1, graph.d_label2);
            gpuErrorcheck(hipDeviceSynchronize());
        }
        subgen.generate(graph, subgraph);
    }
    std::cout << "Number of iterations = " << iterations << std::endl;
    gpuErrorcheck(hipMemcpy(graph.value, graph.d_value, graph.num_nodes * sizeof(uint), hipMemcpyDeviceToHost));
    utilities::SaveResults(out, graph.value, graph.num_nodes);
    return 0;
}

// PageRank with the FP32 residual of the reference's GraphPR: async runs
// each partition until no vertex pushes (pr_async), sync once (pr_kernel)
static int pr_loop(const std::string &path, ull budget, const std::string &out, float acc, bool async)
{
    GraphPR<OutEdge> graph(path, false);
    graph.ReadGraph();
    for (uint v = 0; v < graph.num_nodes; v++) {
        graph.delta[v] = 0.15f;
        graph.value[v] = 0.0f;
    }
    to_device<OutEdge>(graph);
    gpuErrorcheck(hipMemcpy(graph.d_delta, graph.delta, graph.num_nodes * sizeof(float), hipMemcpyHostToDevice));
    Subgraph<OutEdge> subgraph(graph.num_nodes, graph.num_edges, budget, 1);
    SubgraphGenerator<OutEdge> subgen(graph);
    subgen.generate(graph, subgraph, acc);
    Partitioner<OutEdge> partitioner;
    bool finished = false, *d_finished = nullptr;
    gpuErrorcheck(hipMalloc(&d_finished, sizeof(bool)));
    while (subgraph.numActiveNodes > 0) {
        partitioner.partition(subgraph, subgraph.numActiveNodes);
        for (uint i = 0; i < partitioner.numPartitions; i++) {
            fits(partitioner.partitionEdgeSize[i], subgraph.max_partition_size);
            gpuErrorcheck(hipMemcpy(subgraph.d_activeEdgeList, subgraph.activeEdgeList + partitioner.fromEdge[i],
                                    partitioner.partitionEdgeSize[i] * sizeof(OutEdge), hipMemcpyHostToDevice));
            if (!async) {
                pr_kernel((uint)partitioner.partitionNodeSize[i], (uint)partitioner.fromNode[i],
                          partitioner.fromEdge[i], subgraph.d_activeNodes, subgraph.d_activeNodesPointer,
                          subgraph.d_activeEdgeList, graph.d_outDegree, graph.d_value, graph.d_delta, acc);
                gpuErrorcheck(hipDeviceSynchronize());
                continue;
            }
            do {
                finished = true;
                gpuErrorcheck(hipMemcpy(d_finished, &finished, sizeof(bool), hipMemcpyHostToDevice));
                pr_async((uint)partitioner.partitionNodeSize[i], (uint)partitioner.fromNode[i], partitioner.fromEdge[i],
                         subgraph.d_activeNodes, subgraph.d_activeNodesPointer, subgraph.d_activeEdgeList,
                         graph.d_outDegree, graph.d_value, graph.d_delta, d_finished, acc);
                gpuErrorcheck(hipDeviceSynchronize());
                gpuErrorcheck(hipMemcpy(&finished, d_finished, sizeof(bool), hipMemcpyDeviceToHost));
            } while (!finished);
        }
        subgen.generate(graph, subgraph, acc);
    }
    gpuErrorcheck(hipMemcpy(graph.value, graph.d_value, graph.num_nodes * sizeof(float), hipMemcpyDeviceToHost));
    utilities::SaveResults(out, graph.value, graph.num_nodes);
    gpuErrorcheck(hipFree(d_finished));
    return 0;
}

int main(int argc, char **argv)
{
    if (argc != 6) {
        std::cerr << "usage: ref_loop <app> <graph file> <source> <budget bytes> <output file>\n";
        return 2;
    }
    const std::string app = argv[1], path = argv[2], out = argv[5];
    const uint src = (uint)std::strtoul(argv[3], nullptr, 10);
    const ull budget = std::strtoull(argv[4], nullptr, 10);
    const uint INF = DIST_INFINITY;
    if (app == "bfs-async")
        return label_async<OutEdge>(path, src, budget, out, false, INF, 0, false, [](auto... a) { bfs_async(a...); });
    if (app == "sssp-async")
        return label_async<OutEdgeWeighted>(path, src, budget, out, true, INF, 0, false,
                                            [](auto... a) { sssp_async(a...); });
    if (app == "sswp-async")
        return label_async<OutEdgeWeighted>(path, src, budget, out, true, 0, INF, false,
                                            [](auto... a) { sswp_async(a...); });
    if (app == "cc-async")
        return label_async<OutEdge>(path, src, budget, out, false, 0, 0, true, [](auto... a) { cc_async(a...); });
    if (app == "bfs-sync")
        return label_sync<OutEdge>(path, src, budget, out, false, INF, 0, false, [](auto... a) { bfs_kernel(a...); });
    if (app == "sssp-sync")
        return label_sync<OutEdgeWeighted>(path, src, budget, out, true, INF, 0, false,
                                           [](auto... a) { sssp_kernel(a...); });
    if (app == "sswp-sync")
        return label_sync<OutEdgeWeighted>(path, src, budget, out, true, 0, INF, false,
                                           [](auto... a) { sswp_kernel(a...); });
    if (app == "cc-sync")
        return label_sync<OutEdge>(path, src, budget, out, false, 0, 0, true, [](auto... a) { cc_kernel(a...); });
    if (app == "pr-async") return pr_loop(path, budget, out, 1e-6f, true);
    if (app == "pr-sync") return pr_loop(path, budget, out, 1e-6f, false);
    std::cerr << "unknown app " << app << "\n";
    return 2;
}
