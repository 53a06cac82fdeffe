// gpu_kernels.hip -- host launchers with the names and parameters of the
// reference's thirteen kernels (shared/gpu_kernels.cuh:7-131), so an
// application loop written in the reference's order runs unchanged on top
// of this library.  The label kernels reuse the pipeline's frontier
// select-scan and edge-balanced advance (kernels.hip); the PageRank kernels
// and the three label helpers are written here.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>

#include "device.hpp"
#include "kernels.hpp"
#include "subway/gpu_error_check.hpp"
#include "subway/gpu_kernels.hpp"

namespace subway {

using namespace dev;

namespace {

// Scratch the label launchers' frontier and advance need, per device, grown
// on demand: the select-scan descriptors and control block, one worklist
// entry per item, one chunk-map word per 2048 edges, the update counter
// before the launch (for the async kernels' finished flag), and a pinned
// pair of offsets read back to size the rest.
struct LaunchCtx {
    ScanScratch fr;
    Entry *entries = nullptr;
    u64 entryCap = 0;
    u32 *chunkMap = nullptr;
    u64 chunkCap = 0;
    u64 *snap = nullptr;
    u64 *span = nullptr;
    int advGrid = 0;
};
constexpr int kMaxDevices = 64;
std::mutex g_mu;
LaunchCtx g_ctx[kMaxDevices];

LaunchCtx &ctx_for_device()
{
    int dev = 0;
    SUBWAY_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(ERR_ARG, "device index beyond the launcher table");
    LaunchCtx &c = g_ctx[dev];
    if (!c.snap) {
        SUBWAY_CHECK(hipMalloc(&c.snap, sizeof(u64)));
        SUBWAY_CHECK(hipHostMalloc((void **)&c.span, 3 * sizeof(u64), 0));
        c.span[2] = 0;
        hipDeviceProp_t prop;
        SUBWAY_CHECK(hipGetDeviceProperties(&prop, dev));
        c.advGrid = prop.multiProcessorCount * 4;
    }
    return c;
}

void grow(LaunchCtx &c, u64 items, u64 edges)
{
    const u64 chunks = edges / ADV_CHUNK + 2;
    const u64 tiles0 = (items + FR_TILE - 1) / FR_TILE + 1;
    if (tiles0 > c.fr.maxTiles || items + 1 > c.entryCap || chunks > c.chunkCap)
        SUBWAY_CHECK(hipDeviceSynchronize());   // no launch may still use what is replaced
    const u64 tiles = (items + FR_TILE - 1) / FR_TILE + 1;
    if (tiles > c.fr.maxTiles) {
        const u32 spin = c.fr.ctl ? c.fr.spinLimit : SPIN_LIMIT_DEFAULT;
        free_scan_scratch(c.fr);
        alloc_scan_scratch(c.fr, std::max<u64>(tiles, 64));
        c.fr.spinLimit = spin;
    }
    if (items + 1 > c.entryCap) {
        if (c.entries) SUBWAY_CHECK(hipFree(c.entries));
        c.entryCap = std::max<u64>(items + 1, 4096);
        SUBWAY_CHECK(hipMalloc(&c.entries, c.entryCap * sizeof(Entry)));
    }
    if (chunks > c.chunkCap) {
        if (c.chunkMap) SUBWAY_CHECK(hipFree(c.chunkMap));
        c.chunkCap = std::max<u64>(chunks, 1024);
        SUBWAY_CHECK(hipMalloc(&c.chunkMap, c.chunkCap * sizeof(u32)));
    }
}

__global__ void snapshot_kernel(const ScanCtl *ctl, u64 *snap) { *snap = ld_relaxed(&ctl->updates); }

// *finished = false when the launch improved any destination (the async
// kernels' flag, shared/gpu_kernels.cu:345-350)
__global__ void finished_kernel(const ScanCtl *ctl, const u64 *snap, bool *finished)
{
    if (ld_relaxed(&ctl->updates) != *snap) *finished = false;
}

// One launch of a label kernel: the labelled vertices among activeNodes[from
// .. from + numNodes) clear label1 and push their value over the partition's
// edges; improved destinations get label2.
void label_launch(int algo, int fmt, uint numNodes, uint from, ull numPartitionedEdges, const uint *activeNodes,
                  const ull *activeNodesPointer, const void *edgeList, uint *value, bool *finished, bool *label1,
                  bool *label2, hipStream_t st)
{
    if (numNodes == 0) return;
    LaunchCtx &c = ctx_for_device();
    std::lock_guard<std::mutex> lock(g_mu);
    // the partition's edge span sizes the chunk map
    SUBWAY_CHECK(hipMemcpyAsync(&c.span[0], activeNodesPointer + from, sizeof(u64), hipMemcpyDeviceToHost, st));
    SUBWAY_CHECK(hipMemcpyAsync(&c.span[1], activeNodesPointer + from + numNodes, sizeof(u64),
                                hipMemcpyDeviceToHost, st));
    // the error word of the previous launch's bounded waits (0 when none gave up)
    if (c.fr.ctl) SUBWAY_CHECK(hipMemcpyAsync(&c.span[2], &c.fr.ctl->error, sizeof(u32), hipMemcpyDeviceToHost, st));
    SUBWAY_CHECK(hipStreamSynchronize(st));
    if ((u32)c.span[2])
        throw Error(ERR_INTERNAL, std::string("a device wait of the previous label kernel gave up (") +
                                      wait_error_name((u32)c.span[2]) + ")");
    if (c.span[1] < c.span[0]) throw Error(ERR_ARG, "activeNodesPointer is not non-decreasing");
    grow(c, numNodes, c.span[1] - c.span[0]);
    if (finished) hipLaunchKernelGGL(snapshot_kernel, dim3(1), dim3(1), 0, st, c.fr.ctl, c.snap);
    FrontierArgs f{};
    f.algo = algo;
    f.n = numNodes;
    f.nodes = activeNodes + from;
    f.ptr = activeNodesPointer + from;
    f.fromEdge = numPartitionedEdges;
    f.labA = reinterpret_cast<uint8_t *>(label1);
    f.labB = nullptr;
    f.value = value;
    f.entries = c.entries;
    f.chunkMap = c.chunkMap;
    f.first = 1;
    launch_frontier(f, c.fr, st);
    AdvanceArgs a{};
    a.algo = algo;
    a.edgeFmt = fmt;
    a.packCode = 0;
    a.entries = c.entries;
    a.chunkMap = c.chunkMap;
    a.edges = edgeList;
    a.value = value;
    a.labelOut = reinterpret_cast<uint8_t *>(label2);
    a.w0 = 0;
    a.w1 = ~0ull;
    a.shift = 0;
    a.grid = c.advGrid;
    launch_advance(a, c.fr.ctl, st);
    if (finished) hipLaunchKernelGGL(finished_kernel, dim3(1), dim3(1), 0, st, c.fr.ctl, c.snap, finished);
    SUBWAY_CHECK(hipGetLastError());
}

// PageRank push, one wave per active vertex (shared/gpu_kernels.cu:211-254,
// 459-502): a vertex whose residual exceeds acc adds it to its value, sends
// (float)((d / deg) * 0.85) along every out-edge with an FP32 atomic add,
// and subtracts d from its own residual.
__global__ __launch_bounds__(256) void pr_push_kernel(uint numNodes, uint from, u64 base, const uint *activeNodes,
                                                      const ull *anp, const OutEdge *edges, const uint *outDegree,
                                                      float *dist, float *delta, bool *finished, float acc)
{
    const u64 t = ((u64)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    if (t >= numNodes) return;
    const int lane = lane_id();
    const uint id = activeNodes[from + t];
    const uint degree = outDegree[id];
    const float d = delta[id];
    if (!(d > acc)) return;
    if (lane == 0) dist[id] += d;
    if (degree) {
        const float share = (float)((double)(d / (float)degree) * 0.85);
        const u64 e0 = anp[from + t] - base;
        for (u64 i = e0 + lane; i < e0 + degree; i += WAVE) unsafeAtomicAdd(&delta[edges[i].end], share);
        if (finished && lane == 0) *finished = false;
    }
    if (lane == 0) unsafeAtomicAdd(&delta[id], -d);
}

void pr_launch(uint numNodes, uint from, ull base, const uint *activeNodes, const ull *anp, const OutEdge *edges,
               const uint *outDegree, float *dist, float *delta, bool *finished, float acc, hipStream_t st)
{
    if (numNodes == 0) return;
    const u64 threads = (u64)numNodes * WAVE;
    hipLaunchKernelGGL(pr_push_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, numNodes, from,
                       (u64)base, activeNodes, anp, edges, outDegree, dist, delta, finished, acc);
    SUBWAY_CHECK(hipGetLastError());
}

// 0: clearLabel, 1: mixLabels, 2: moveUpLabels (shared/gpu_kernels.cu:506-534)
__global__ void labels_kernel(int op, const uint *activeNodes, bool *l1, bool *l2, uint size, uint from)
{
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= size) return;
    const uint v = activeNodes[from + i];
    if (op == 0) {
        l1[v] = false;
    } else if (op == 1) {
        l1[v] = l1[v] || l2[v];
        l2[v] = false;
    } else {
        l1[v] = l2[v];
        l2[v] = false;
    }
}

void labels_launch(int op, const uint *activeNodes, bool *l1, bool *l2, uint size, uint from, hipStream_t st)
{
    if (size == 0) return;
    hipLaunchKernelGGL(labels_kernel, dim3((size + 255) / 256), dim3(256), 0, st, op, activeNodes, l1, l2, size,
                       from);
    SUBWAY_CHECK(hipGetLastError());
}

// The reference's error convention for these entry points: message on
// stderr, exit with the code (shared/gpu_error_check.cuh:8-16).
template <class F>
void or_exit(F &&f)
{
    try {
        f();
    } catch (const Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        std::exit(e.code > 0 ? e.code : 1);
    }
}

}  // namespace
}  // namespace subway

using namespace subway;

void bfs_kernel(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
                OutEdge *edgeList, uint *outDegree, uint *value, bool *label1, bool *label2, hipStream_t stream)
{
    (void)outDegree;
    or_exit([&] {
        label_launch(ALGO_BFS, EDGE_U32, numNodes, from, numPartitionedEdges, activeNodes, activeNodesPointer,
                     edgeList, value, nullptr, label1, label2, stream);
    });
}

void cc_kernel(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
               OutEdge *edgeList, uint *outDegree, uint *dist, bool *label1, bool *label2, hipStream_t stream)
{
    (void)outDegree;
    or_exit([&] {
        label_launch(ALGO_CC, EDGE_U32, numNodes, from, numPartitionedEdges, activeNodes, activeNodesPointer,
                     edgeList, dist, nullptr, label1, label2, stream);
    });
}

void sssp_kernel(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
                 OutEdgeWeighted *edgeList, uint *outDegree, uint *dist, bool *label1, bool *label2,
                 hipStream_t stream)
{
    (void)outDegree;
    or_exit([&] {
        label_launch(ALGO_SSSP, EDGE_W64, numNodes, from, numPartitionedEdges, activeNodes, activeNodesPointer,
                     edgeList, dist, nullptr, label1, label2, stream);
    });
}

void sswp_kernel(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
                 OutEdgeWeighted *edgeList, uint *outDegree, uint *dist, bool *label1, bool *label2,
                 hipStream_t stream)
{
    (void)outDegree;
    or_exit([&] {
        label_launch(ALGO_SSWP, EDGE_W64, numNodes, from, numPartitionedEdges, activeNodes, activeNodesPointer,
                     edgeList, dist, nullptr, label1, label2, stream);
    });
}

void pr_kernel(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
               OutEdge *edgeList, uint *outDegree, float *dist, float *delta, float acc, hipStream_t stream)
{
    or_exit([&] {
        pr_launch(numNodes, from, numPartitionedEdges, activeNodes, activeNodesPointer, edgeList, outDegree, dist,
                  delta, nullptr, acc, stream);
    });
}

void bfs_async(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
               OutEdge *edgeList, uint *outDegree, uint *dist, bool *finished, bool *label1, bool *label2,
               hipStream_t stream)
{
    (void)outDegree;
    or_exit([&] {
        label_launch(ALGO_BFS, EDGE_U32, numNodes, from, numPartitionedEdges, activeNodes, activeNodesPointer,
                     edgeList, dist, finished, label1, label2, stream);
    });
}

void sssp_async(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
                OutEdgeWeighted *edgeList, uint *outDegree, uint *dist, bool *finished, bool *label1, bool *label2,
                hipStream_t stream)
{
    (void)outDegree;
    or_exit([&] {
        label_launch(ALGO_SSSP, EDGE_W64, numNodes, from, numPartitionedEdges, activeNodes, activeNodesPointer,
                     edgeList, dist, finished, label1, label2, stream);
    });
}

void sswp_async(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
                OutEdgeWeighted *edgeList, uint *outDegree, uint *dist, bool *finished, bool *label1, bool *label2,
                hipStream_t stream)
{
    (void)outDegree;
    or_exit([&] {
        label_launch(ALGO_SSWP, EDGE_W64, numNodes, from, numPartitionedEdges, activeNodes, activeNodesPointer,
                     edgeList, dist, finished, label1, label2, stream);
    });
}

void cc_async(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
              OutEdge *edgeList, uint *outDegree, uint *dist, bool *finished, bool *label1, bool *label2,
              hipStream_t stream)
{
    (void)outDegree;
    or_exit([&] {
        label_launch(ALGO_CC, EDGE_U32, numNodes, from, numPartitionedEdges, activeNodes, activeNodesPointer,
                     edgeList, dist, finished, label1, label2, stream);
    });
}

void pr_async(uint numNodes, uint from, ull numPartitionedEdges, uint *activeNodes, ull *activeNodesPointer,
              OutEdge *edgeList, uint *outDegree, float *dist, float *delta, bool *finished, float acc,
              hipStream_t stream)
{
    or_exit([&] {
        pr_launch(numNodes, from, numPartitionedEdges, activeNodes, activeNodesPointer, edgeList, outDegree, dist,
                  delta, finished, acc, stream);
    });
}

void clearLabel(uint *activeNodes, bool *label, uint size, uint from, hipStream_t stream)
{
    or_exit([&] { labels_launch(0, activeNodes, label, nullptr, size, from, stream); });
}

void mixLabels(uint *activeNodes, bool *label1, bool *label2, uint size, uint from, hipStream_t stream)
{
    or_exit([&] { labels_launch(1, activeNodes, label1, label2, size, from, stream); });
}

void moveUpLabels(uint *activeNodes, bool *label1, bool *label2, uint size, uint from, hipStream_t stream)
{
    or_exit([&] { labels_launch(2, activeNodes, label1, label2, size, from, stream); });
}
