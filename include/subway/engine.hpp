// engine.hpp -- the global-iteration loop of Subway's applications
// (reference: the main() loops of subway/{bfs,cc,sssp,sswp,pr}-{sync,async}.cu,
// e.g. subway/sssp-async.cu:77-122 and subway/bfs-sync.cu:57-89).
//
// One call runs an algorithm from its initial state to the fixpoint:
//   generate (device) -> partition (device) -> for each partition: transfer
//   on the copy stream into one of two device buffers while the previous
//   partition computes -> frontier + advance rounds (one round in sync mode,
//   rounds until the partition converges in async mode) -> generate ...
// Results are left in the graph's host value array.
#ifndef SUBWAY_ENGINE_HPP
#define SUBWAY_ENGINE_HPP

#include <string>
#include <vector>

#include "globals.hpp"
#include "graph.hpp"

namespace subway {

enum AlgoId { BFS = 0, CC = 1, SSSP = 2, SSWP = 3, PR = 4 };
enum ModeId { SYNC = 0, ASYNC = 1 };
// How a compacted partition reaches the device.
enum TransferId {
    TRANSFER_AUTO = 0,      // DMA when the partition is a contiguous CSR slice, else GPU gather
    TRANSFER_ZEROCOPY = 1,  // GPU gather from mapped pinned host memory
    TRANSFER_HOST = 2,      // host threads gather into pinned staging, then DMA
};
enum PolicyId { POLICY_PLAIN = 0, POLICY_LEARNED = 1 };

struct RunOptions {
    int algo = BFS;
    int mode = SYNC;
    uint source = 0;
    ull budgetBytes = 0;    // device edge budget in bytes; 0 = reference sizing
    int numBuffers = 2;     // 1 = no transfer/compute overlap
    float acc = 0.01f;      // PR residual threshold (pr-*.cu:35)
    float initPR = 0.15f;   // PR initial delta (pr-*.cu:34)
    int transfer = TRANSFER_AUTO;
    int policy = POLICY_PLAIN;
    int verbose = 0;        // 1: print the reference's per-partition lines
    int advanceGrid = 0;    // persistent workgroups of the advance kernel; 0 = auto
    int maxRoundBatch = 32; // async: rounds issued between convergence checks
    int profile = 0;        // 1: time every kernel launch and transfer with HIP events
};

struct IterRecord {
    uint iteration;
    ull activeNodes;
    ull activeEdges;
    uint partitions;
    ull h2dBytes;
    ull rounds;
    ull updates;
    double msIteration;
};

struct RunStats {
    double msTotal = 0;        // first generate .. values on the host
    double msTransfer = 0;     // copy-stream busy time (DMA + gather kernels), profile only
    double msAdvance = 0;      // advance kernel time, profile only
    double msFrontier = 0;     // frontier kernel time, profile only
    double msGenerate = 0;     // generator + partition kernels, profile only
    ull advanceLaunches = 0;
    ull frontierLaunches = 0;
    ull transferOps = 0;
    ull workEdges = 0;         // edges relaxed: sum of frontier work over all rounds
    ull workEntries = 0;       // vertices that pushed: sum over all rounds
    ull h2dBytes = 0;          // edge bytes moved host -> device
    ull h2dDmaBytes = 0;       // ... of which by DMA
    ull h2dGatherBytes = 0;    // ... of which by GPU gather reads
    uint globalIterations = 0;
    ull innerRounds = 0;       // frontier launches that selected work
    ull partitions = 0;
    ull updates = 0;           // successful relaxations / PR pushes
    std::vector<IterRecord> trace;
};

template <class E>
RunStats runLabel(Graph<E> &graph, const RunOptions &opt);
template <class E>
RunStats runPageRank(GraphPR<E> &graph, const RunOptions &opt);

const char *algo_name(int algo);

}  // namespace subway

#endif  // SUBWAY_ENGINE_HPP
