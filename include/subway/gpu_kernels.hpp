// gpu_kernels.hpp -- the reference's thirteen kernels as host launchers
// (drop-in for shared/gpu_kernels.cuh:7-131).
//
// The reference launches each kernel itself, one thread per active vertex
// (`sssp_async<<<partitionNodeSize/512 + 1, 512>>>(...)`, subway/sssp-async.cu:99-110).
// Here every name is an ordinary function with the reference's parameters in
// the reference's order, plus an optional stream; launch geometry is chosen
// inside:
//   * the eight label kernels (bfs/cc/sssp/sswp, sync and async) select the
//     labelled vertices of the partition with a wave64 select-scan and relax
//     their edges edge-balanced (work chunks of 2048 edges claimed by
//     persistent workgroups), so a hub vertex is spread over many workgroups
//     instead of one thread;
//   * the two PageRank kernels push from one wave per active vertex (lanes
//     stride its edges, FP32 atomic adds of the reference's share);
//   * clearLabel / mixLabels / moveUpLabels are one lane per item.
// The update rules are the reference's (shared/gpu_kernels.cu:9-534): the
// candidate value[u] + 1 / value[u] / value[u] + w / min(value[u], w), a
// 32-bit atomicMin (atomicMax for SSWP) when it improves the destination,
// label2 set for improved destinations, label1 of the pushing vertex
// cleared, and for the async kernels *finished = false when any destination
// improved.  A launcher returns once the work is enqueued on `stream`; like
// the reference's kernels the results are visible after a synchronisation.
// Each call reads two offsets of activeNodesPointer back to the host to size
// its scratch, so it synchronises `stream` once before launching.
//
// Differences from the reference a caller can see: activeNodesPointer and
// numPartitionedEdges are 64-bit (graphs of 2^32 edges and more);
// the edge range of a vertex is [activeNodesPointer[from + t] -
// numPartitionedEdges, + outDegree[id]) as in the reference.  Errors follow
// the reference's convention (gpuErrorcheck: message and exit).
#ifndef SUBWAY_GPU_KERNELS_HPP
#define SUBWAY_GPU_KERNELS_HPP

#include <hip/hip_runtime.h>

#include "globals.hpp"

// shared/gpu_kernels.cuh:7-17 (body shared/gpu_kernels.cu:9-58)
void bfs_kernel(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
                OutEdge *edgeList, uint *outDegree, uint *value, bool *label1, bool *label2,
                hipStream_t stream = nullptr);
// shared/gpu_kernels.cuh:19-29 (gpu_kernels.cu:60-108)
void cc_kernel(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
               OutEdge *edgeList, uint *outDegree, uint *dist, bool *label1, bool *label2,
               hipStream_t stream = nullptr);
// shared/gpu_kernels.cuh:31-41 (gpu_kernels.cu:111-159)
void sssp_kernel(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
                 OutEdgeWeighted *edgeList, uint *outDegree, uint *dist, bool *label1, bool *label2,
                 hipStream_t stream = nullptr);
// shared/gpu_kernels.cuh:43-53 (gpu_kernels.cu:161-209)
void sswp_kernel(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
                 OutEdgeWeighted *edgeList, uint *outDegree, uint *dist, bool *label1, bool *label2,
                 hipStream_t stream = nullptr);
// shared/gpu_kernels.cuh:55-65 (gpu_kernels.cu:211-254)
void pr_kernel(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
               OutEdge *edgeList, uint *outDegree, float *dist, float *delta, float acc,
               hipStream_t stream = nullptr);
// shared/gpu_kernels.cuh:67-77 (gpu_kernels.cu:257-306)
void bfs_async(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
               OutEdge *edgeList, uint *outDegree, uint *dist, bool *finished, bool *label1, bool *label2,
               hipStream_t stream = nullptr);
// shared/gpu_kernels.cuh:79-89 (gpu_kernels.cu:308-356)
void sssp_async(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
                OutEdgeWeighted *edgeList, uint *outDegree, uint *dist, bool *finished, bool *label1, bool *label2,
                hipStream_t stream = nullptr);
// shared/gpu_kernels.cuh:91-101 (gpu_kernels.cu:358-406)
void sswp_async(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
                OutEdgeWeighted *edgeList, uint *outDegree, uint *dist, bool *finished, bool *label1, bool *label2,
                hipStream_t stream = nullptr);
// shared/gpu_kernels.cuh:103-113 (gpu_kernels.cu:408-457)
void cc_async(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
              OutEdge *edgeList, uint *outDegree, uint *dist, bool *finished, bool *label1, bool *label2,
              hipStream_t stream = nullptr);
// shared/gpu_kernels.cuh:115-125 (gpu_kernels.cu:459-502)
void pr_async(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
              OutEdge *edgeList, uint *outDegree, float *dist, float *delta, bool *finished, float acc,
              hipStream_t stream = nullptr);
// shared/gpu_kernels.cuh:127-131 (gpu_kernels.cu:506-534): label[activeNodes[from + t]] = false;
// label1 |= label2, label2 = false; label1 = label2, label2 = false
void clearLabel(uint *activeNodes, bool *label, uint size, uint from, hipStream_t stream = nullptr);
void mixLabels(uint *activeNodes, bool *label1, bool *label2, uint size, uint from, hipStream_t stream = nullptr);
void moveUpLabels(uint *activeNodes, bool *label1, bool *label2, uint size, uint from, hipStream_t stream = nullptr);

#endif  // SUBWAY_GPU_KERNELS_HPP
