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

#include <memory>
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
    int persistent = 1;     // async: all rounds of a partition in one persistent launch
    int roundsGrid = 0;     // workgroups of the persistent kernel; 0 (and the cap) = one below full occupancy per CU
    int maxRounds = 4096;     // async rounds per partition visit (labels left set resume next iteration)
    int batch = 0;          // 1: batched k-source queries (runBatch), BFS/SSSP/SSWP async
    int batchUnroll = 0;    // batched advance: edges in flight per wave (1 .. 32, powers of 2); 0 = 8
    float theta = 0.5f;     // learned policy without a model: whole-range density threshold
    uint32_t spinLimit = 0; // polls before a device wait gives up (the run then fails); 0 = default
    uint32_t joinWindow = 0;   // persistent launches: membership window in 100 MHz ticks; 0 = 50000 (500 us)
    int unpackedEdges = 0;  // 1: weighted edges stay 8 bytes; 0: one 32-bit word when ids and weights fit
    int noResident = 0;     // 1: move active edges every iteration even when the whole graph fits one buffer
    uint residentMinBytes = 0;   // learned schedule: least active edge bytes for an in-place visit (0: 8 MiB)
    int pushOnly = 0;       // BFS on a resident graph: 1 = top-down steps only (no direction switching)
    int noBlocks = 0;       // 1: whole id ranges move as plain edges even when the graph holds its block layout
    int earlyReadback = 0;  // 0: early copy of the results on large graphs; -1: never; N > 0: always, patch list N
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
    // learned policy: the iteration's features and network heads, and the
    // schedule applied (whole-range density threshold, inner-round cap)
    float features[12] = {};
    float heads[3] = {};
    float theta = 0.f;
    uint roundsCap = 0;
    uint onPlain = 0;          // 1: the plain policy scheduled this iteration (fallback or no network)
};

struct RunStats {
    double msTotal = 0;        // first generate .. values on the host
    double msTransfer = 0;     // copy-stream busy time (DMA + gather kernels), profile only
    double msAdvance = 0;      // advance kernel time, profile only
    double msFrontier = 0;     // frontier kernel time, profile only
    double msGenerate = 0;     // generator + partition kernels, profile only
    double msRounds = 0;       // persistent rounds kernel time, profile only
    ull roundsLaunches = 0;
    ull advanceLaunches = 0;
    ull frontierLaunches = 0;
    ull transferOps = 0;
    ull workEdges = 0;         // edges relaxed: sum of frontier work over all rounds
    ull workEntries = 0;       // vertices that pushed: sum over all rounds
    ull workItems = 0;         // partition items scanned: sum over all rounds
    ull laneWork = 0;          // batched runs: (edge, active source) pairs relaxed over all rounds
    ull laneEntries = 0;       // batched runs: active sources of the pushing vertices over all rounds
    // the part of workEdges, workEntries and updates done by stand-alone
    // frontier + advance launches (round cap of one, launch-per-round mode,
    // oversized vertices) rather than persistent rounds launches
    ull pairWorkEdges = 0;
    ull pairWorkEntries = 0;
    ull pairUpdates = 0;
    uint residentVisits = 0;   // learned schedule: units visited in place on a slice a buffer still held
    uint residentOverlaps = 0; // ... resident slices skipped because an overlapping one was visited
    uint pullRounds = 0;       // direction-optimizing BFS: bottom-up steps
    ull pullFound = 0;         // ... vertices they found
    uint mispredictions = 0;   // learned policy: iterations whose active edges missed the prediction by > 2x
    uint fallbackIteration = 0;   // learned policy: first iteration scheduled by the plain fallback (0: none)
    uint edgeBytes = 0;        // bytes per edge in transfers and device buffers (4 when packed)
    uint minMembers = 0;       // fewest member workgroups of a persistent launch (0: none ran)
    ull h2dBytes = 0;          // edge bytes moved host -> device
    ull h2dDmaBytes = 0;       // ... of which by DMA
    ull h2dGatherBytes = 0;    // ... of which by GPU gather reads
    ull h2dBlockedBytes = 0;   // ... of which in the block layout
    double msRoundsBlocked = 0;   // persistent rounds launches over block units: time (profile only), count,
    ull roundsLaunchesBlocked = 0;  // and their share of workEdges, workEntries and updates
    ull blkWorkEdges = 0;
    ull blkWorkEntries = 0;
    ull blkUpdates = 0;
    ull blockedEdges = 0;      // edges of the units moved in the block layout
    uint blockedUnits = 0;     // units moved in the block layout
    uint earlyReadback = 0;    // 0: results copied at the end; 1: copied after the first iteration, then the
                               // changed values patched; 2: ... then copied again in full (too many changed)
    ull valuesPatched = 0;     // values that changed after the early copy
    uint globalIterations = 0;
    ull innerRounds = 0;       // frontier launches that selected work
    ull partitions = 0;
    ull updates = 0;           // successful relaxations / PR pushes
    std::vector<IterRecord> trace;
};

// A graph's device state and all pipeline buffers, built once and reused
// for many runs (e.g. queries from different sources).
class Session {
public:
    virtual ~Session() = default;
    virtual RunStats run(uint source) = 0;   // the source is ignored by CC and PR
    virtual void values(void *out) = 0;      // result of the last run (u32 or float per vertex)
    virtual const void *valuesPtr() = 0;     // the graph's host value array (last run's results)
    virtual ull numNodes() const = 0;
    // Batched sessions (RunOptions::batch): k <= 64 sources per run; rows()
    // copies the n x 64 vertex-major values of the last run (lane i = sources[i]).
    virtual RunStats runBatch(const uint *sources, int k) = 0;
    virtual void rows(uint *out) = 0;
    // Autotuner of the batched advance: times complete runs (exact results)
    // for each candidate wave assignment and keeps the fastest.  Returns the
    // chosen edges-in-flight per wave; `timesMs` (may be null) gets 4 times.
    virtual int tuneBatch(const uint *sources, int k, double *timesMs) = 0;
    // Scheduler network parameters (SUBWAY_MLP_PARAMS floats) for the learned
    // policy; null removes them (fixed theta / inner-round cap instead).
    virtual void setPolicy(const float *params) = 0;
};

template <class E>
std::unique_ptr<Session> make_session(std::shared_ptr<HostCsr<E>> csr, const RunOptions &opt);

template <class E>
RunStats runLabel(Graph<E> &graph, const RunOptions &opt);
template <class E>
RunStats runPageRank(GraphPR<E> &graph, const RunOptions &opt);

const char *algo_name(int algo);

// Destroys the compute/copy stream pair every engine of the process shares
// (after draining it).  Call it only when no session or engine is alive, e.g.
// at the end of a program; a later engine makes a new pair.  Also runs by
// itself at process exit, before the HIP runtime's teardown.
void release_shared_streams();

}  // namespace subway

#endif  // SUBWAY_ENGINE_HPP
