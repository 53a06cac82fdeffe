// kernels.hpp -- host-side launchers of the gfx950 kernels that make up the
// out-of-memory pipeline.  Every launcher enqueues on the given stream and
// returns immediately; shapes are validated on the host before launch.
#ifndef SUBWAY_KERNELS_HPP
#define SUBWAY_KERNELS_HPP

#include <hip/hip_runtime.h>

#include <cstdint>

#include "scan_types.hpp"
#include "subway/globals.hpp"

namespace subway {

using dev::ScanCtl;
using dev::TileDesc;

enum Algo { ALGO_BFS = 0, ALGO_CC = 1, ALGO_SSSP = 2, ALGO_SSWP = 3, ALGO_PR = 4 };

// Select-scan tiling (generator and frontier kernels).
constexpr int SCAN_BLOCK = 256;
constexpr int GEN_IPT = 16;                            // vertices per thread
constexpr int GEN_TILE = SCAN_BLOCK * GEN_IPT;         // 4096 vertices per tile
constexpr int GEN_SUB = 4;                              // tiles per generator chunk
constexpr int GEN_CHUNK = GEN_TILE * GEN_SUB;           // 16384 vertices per workgroup of the generator
constexpr int FR_IPT = 8;
constexpr int FR_TILE = SCAN_BLOCK * FR_IPT;           // 2048 active-list items per tile
// Advance kernel: work space is cut into chunks of ADV_CHUNK edges.
constexpr int ADV_BLOCK = 256;
constexpr int ADV_CHUNK = 2048;
// Gather kernel: destination chunk per workgroup.
constexpr int GATHER_BLOCK = 256;
constexpr int GATHER_CHUNK = 4096;

// One worklist entry per selected vertex: where its edges sit in the device
// buffer, where its edges start in work space, and the value it pushes
// (uint32 value, or float bits of the PageRank share).
// Device formats of the edges in a partition buffer.  The unpacked formats'
// values are their sizes in bytes.  EDGE_PACKED stores a weighted edge in one
// word, destination | (weight - base) << shift, for graphs whose vertex ids
// fit `shift` bits and whose weight span fits the remaining 32 - shift bits;
// kernels take the decode word packCode = shift | base << 5.  EDGE_SPLIT
// keeps a buffer of capacity C edges as C u32 destinations followed, at byte
// 4 C, by C u8 weights (5 bytes per edge); kernels take packCode = C.
// EDGE_BLOCKS holds the block-compressed rows of an id range (blocks.hpp);
// an entry's edge_local is then its row's word offset in the buffer, and
// kernels take packCode = the weight code (bits | base << 6).
enum EdgeFormat : int {
    EDGE_U32 = 4,      // OutEdge
    EDGE_W64 = 8,      // OutEdgeWeighted
    EDGE_PACKED = 5,   // PackedEdge
    EDGE_SPLIT = 6,    // SplitEdge
    EDGE_BLOCKS = 7,   // BlockEdge
};
struct PackedEdge {
    uint32_t bits;
};
struct SplitEdge {
    uint32_t dst;      // element of the destination array; the weights follow the array
};
struct BlockEdge {
    uint32_t word;     // word of the block layout
};

struct Entry {
    uint32_t edge_local;
    uint32_t prefix;
    uint32_t val;
    uint32_t degree;   // out-degree of the vertex (the block layout decodes from it)
};

// One partition row as produced by the device partitioner.
struct PartRow {
    ull fromNode;
    ull nodeCount;
    ull fromEdge;
    ull edgeCount;
    ull firstId;
    ull lastId;
};

// Scratch for one select-scan kernel family.
struct ScanScratch {
    ScanCtl *ctl = nullptr;
    TileDesc td{};
    ull maxTiles = 0;
    // Launch epochs handed out since the descriptors were last cleared.  A
    // descriptor word holds 22 epoch bits, so the descriptors are cleared
    // before that many epochs have passed (a stale word could otherwise
    // alias the running epoch).
    ull epochsIssued = 0;
    uint32_t spinLimit = dev::SPIN_LIMIT_DEFAULT;   // poll limit of the scans' bounded waits
};

// Error word of the scratch's control block (0: no wait gave up); copies it
// to the host and synchronises the stream.
uint32_t scan_wait_error(const ScanScratch &s, hipStream_t st);
const char *wait_error_name(uint32_t code);

void alloc_scan_scratch(ScanScratch &s, ull maxTiles);
void free_scan_scratch(ScanScratch &s);

// ---- subgraph generator (flags -> scan -> compaction, one pass)
// Exclusive scan of the block-row word counts of the active list (n items,
// ids), into wpre[0 .. n] (uses the generator's scan scratch).
void launch_row_words(ull n, const uint32_t *ids, const ull *rowPtr, ull *wpre, ScanScratch &s, hipStream_t st);
void launch_generate_labels(ull n, const uint8_t *l1, const uint8_t *l2,
                            const uint32_t *outDegree, uint32_t *activeNodes,
                            ull *activeNodesPointer, ScanScratch &s, hipStream_t st);
void launch_generate_pr(ull n, const float *delta, float acc,
                        const uint32_t *outDegree, uint32_t *activeNodes,
                        ull *activeNodesPointer, ScanScratch &s, hipStream_t st);
void launch_generate_pr(ull n, const double *delta, double acc,
                        const uint32_t *outDegree, uint32_t *activeNodes,
                        ull *activeNodesPointer, ScanScratch &s, hipStream_t st);

// ---- partitioner: greedy contiguous ranges of the active list
// Writes at most maxRows rows starting at active index `start`; out[0] =
// rows written, out[1] = next start, out[2] = active count, out[3] = active
// edge total.
// The active count is read from dNumActive when it is non-null (so the
// launch can follow the generator without a host round trip).
void launch_partition(const ull *activeNodesPointer, const uint32_t *activeNodes,
                      const ull *dNumActive, ull numActive, ull budgetEdges,
                      ull start, uint32_t maxRows, PartRow *rows, ull *out,
                      hipStream_t st);

// ---- per static id range: active-list slice [lo, hi), its active edges and
// the active-edge offset at lo (out: 4 x u64 per range)
// Scheduler features and heads of one iteration (the policy kernel's output).
struct PolicyOut {
    float feat[12];
    float head[3];
    float pad;
};
// Device state the policy kernel carries from iteration to iteration.
struct PolicyState {
    ull prevUpdates;
};
struct PolicyArgs {
    const uint32_t *activeNodes;
    const ull *anp;                  // activeNodesPointer
    const ull *totals;               // generator [count, total, error]
    ull *hist;                       // generator degree histogram (cleared here)
    const ull *updatesNow;           // cumulative successful updates
    PolicyState *state;
    const uint32_t *bounds;          // P + 1 id-range bounds
    int P;
    const ull *nodePointer;          // device CSR offsets
    const float *params;             // MLP parameters or null (features only)
    ull V, E, bufEdges;
    uint32_t iteration;
    uint32_t pageRank;               // 1: the run is PageRank (a feature of its own)
    PolicyOut *out;
    ull *rstats;                     // P x {lo, hi, active edges, first offset}
};
void launch_policy(const PolicyArgs &a, hipStream_t st);

// ---- zero-copy gather of a partition's adjacency from mapped host memory
void launch_gather(int edgeBytes, const uint32_t *activeNodes, const ull *activeNodesPointer,
                   ull kbeg, ull kend, const ull *nodePointer,
                   const void *hostEdges, void *dst, ull edgeBase, ull edgeCount,
                   hipStream_t st);

// ---- per-round frontier (filter) kernel
// Label algorithms: an item is selected when labA (or labB) is set; both are
// cleared.  PR: selected when delta > acc; the running sum absorbs delta.
struct FrontierArgs {
    int algo;
    ull n;                      // items in the partition
    const uint32_t *nodes;           // activeNodes + fromNode
    const ull *ptr;             // activeNodesPointer + fromNode
    ull fromEdge;               // subtracted from ptr to get buffer offsets
    uint8_t *labA;
    uint8_t *labB;                   // may be null
    uint32_t *value;                 // u32 value (label algos)
    double *delta;                   // PR only: residual
    double *sum;                     // PR only: running value sum
    double acc;                      // PR only
    Entry *entries;
    uint32_t *chunkMap;
    int first;                       // 1: ignore the stop flag of the previous launch
    uint32_t idBase;                 // nodes == null: the items are ids idBase .. idBase + n - 1
    const ull *blk;                  // EDGE_BLOCKS buffers: row word offsets of the graph (else null)
    ull blkBase;                     // ... of the buffer's first row
    uint32_t blkByItem;              // 1: blk is indexed by item (rows gathered in item order), 0: by vertex id
};
void launch_frontier(const FrontierArgs &a, ScanScratch &s, hipStream_t st);

// ---- per-round advance (edge relaxation) kernel
struct AdvanceArgs {
    int algo;
    int edgeFmt;                     // EdgeFormat of the buffer
    uint32_t packCode;              // EDGE_PACKED: shift | weight base << 5
    const Entry *entries;
    const uint32_t *chunkMap;
    const void *edges;               // device edge buffer
    uint32_t *value;                 // u32 values; PR: unused
    uint8_t *labelOut;               // label algos
    double *delta;                   // PR residual
    ull w0, w1;                 // work window [w0, w1); w1 = ~0 for all
    ull shift;                  // subtracted from edge indices
    int grid;                        // persistent workgroups
    uint32_t nv;                     // vertices of the graph (block decode checks destinations against it)
    uint32_t wlim;                   // EDGE_BLOCKS: words of the buffer (the decode reads none at or past it)
};
void launch_advance(const AdvanceArgs &a, ScanCtl *ctl, hipStream_t st);

// ---- persistent async rounds (one launch per partition)
struct RoundsLaunch {
    ull n;
    const uint32_t *nodes;
    const ull *ptr;
    ull fromEdge;
    uint8_t *l1, *l2;
    uint32_t *value;
    double *delta;         // PR residual
    double *sum;           // PR running value sum
    double acc;
    Entry *entries;
    uint32_t *chunkMap;
    const void *edges;
    ScanScratch *scan;     // frontier scan scratch (its epoch budget is charged)
    dev::PersistCtl *pc;
    uint32_t maxRounds;
    uint32_t roundBase;
    uint32_t idBase;       // nodes == null: a whole id range starting here
    uint32_t joinTicks;    // membership window of the launch (100 MHz ticks); 0 = 500 us
    uint32_t packCode;    // EDGE_PACKED: shift | weight base << 5
    const ull *blk;        // EDGE_BLOCKS buffers: row word offsets of the graph (else null)
    ull blkBase;           // ... of the buffer's first row
    uint32_t blkByItem;    // 1: blk is indexed by item (rows gathered in item order), 0: by vertex id
    uint32_t nv;           // vertices of the graph (block decode checks destinations against it)
    uint32_t wlim;         // EDGE_BLOCKS: words of the buffer (the decode reads none at or past it)
};
// Largest co-resident grid for the rounds kernel, one workgroup slot per CU
// below full occupancy.
int rounds_max_grid(int algo, int edgeFmt);
void launch_rounds(int algo, int edgeFmt, const RoundsLaunch &l, int grid, hipStream_t st);

// ---- batched k-source runs (k <= 64): dist[V][64], 64-bit masks
constexpr int BATCH_LANES = 64;
constexpr int BATCH_CHUNK_EDGES = 256;   // work-space chunk of the batched advance
void launch_generate_masks(ull n, const ull *m1, const ull *m2, const uint32_t *outDegree, uint32_t *activeNodes,
                           ull *activeNodesPointer, ScanScratch &s, hipStream_t st);
void launch_init_batch_state(int algo, ull n, uint32_t *dist, ull *m1, ull *m2, int allActive, hipStream_t st);
void launch_set_batch_sources(int algo, const uint32_t *sources, int k, uint32_t *dist, ull *m1, hipStream_t st);
struct BatchLaunch {
    ull n;
    const uint32_t *nodes;
    const ull *ptr;
    ull fromEdge;
    ull *m1, *m2;
    uint32_t *dist;
    Entry *entries;
    ull *entryMask;
    uint32_t *chunkMap;
    const void *edges;
    ScanScratch *scan;
    dev::PersistCtl *pc;
    uint32_t maxRounds;
    uint32_t idBase;
    uint32_t joinTicks;    // membership window of the launch (100 MHz ticks); 0 = 500 us
    uint32_t packCode;
};
int batch_max_grid(int algo, int edgeFmt, int unroll);
void launch_frontier_masks(ull n, const uint32_t *nodes, const ull *ptr, ull fromEdge, ull *labA, ull *labB,
                           Entry *entries, ull *entryMask, uint32_t *chunkMap, ScanScratch &s, hipStream_t st,
                           uint32_t idBase = 0);
// Batched advance over the work window [w0, w1) of the last frontier (a
// vertex larger than a buffer, streamed chunk by chunk).
void launch_advance_batch_window(int algo, int edgeFmt, uint32_t packCode, const Entry *entries,
                                 const ull *entryMask, const uint32_t *chunkMap, const void *edges, uint32_t *dist,
                                 ull *maskOut, ull w0, ull w1, ull shift, ScanCtl *ctl, int grid, hipStream_t st);
void launch_rounds_batch(int algo, int edgeFmt, int unroll, const BatchLaunch &l, int grid, hipStream_t st);

// Initial vertex state (values and both label arrays / PR value and delta).
// ids (may be null): CC's initial value of vertex i, the original id of a
// renumbered graph's vertex i.
void launch_init_label_state(int algo, int async, ull n, uint32_t *value, uint8_t *l1, uint8_t *l2,
                             uint32_t source, const uint32_t *ids, hipStream_t st);
// out[i * width + c] = in[perm[v0 + i] * width + c] for i < count: values (or
// batched rows) of a renumbered graph in the original order.
void launch_unpermute(const uint32_t *in, const uint32_t *perm, ull v0, ull count, uint32_t width, uint32_t *out,
                      hipStream_t st);
// n words into mapped pinned host memory (dstMapped: its device address),
// stored by `grid` workgroups; src and dstMapped 16-byte aligned.
void launch_store_host(const uint32_t *src, uint32_t *dstMapped, ull n, int grid, hipStream_t st);
// The vertices v whose value in[v] differs from snap[v]: (name[v], value)
// pairs (name null: v) into list (the first cap of them), their number
// added to *count (which the caller zeroes).
void launch_value_changes(const uint32_t *in, const uint32_t *name, const uint32_t *snap, ull n, uint32_t *count,
                          uint2 *list, uint32_t cap, hipStream_t st);
// PR: the running sum and the residual are FP64 (an FP32 residual at a hub
// rounds away shares below half its ulp).
void launch_init_pr_state(ull n, double *sum, double *delta, double init, hipStream_t st);
// PR result: FP32 values from the running sums, and (deltaOut non-null) the
// FP32 copy of the final FP64 residual
void launch_finish_pr(ull n, float *value, const double *sum, float *deltaOut, const double *resid,
                      hipStream_t st);

// Bottom-up BFS step over every vertex: unvisited vertices with an
// in-neighbour at level `cur` get level cur + 1 and labOut set; the number
// found and their out-edges are added to ctl->pull_found / pull_edges.
void launch_bfs_pull(ull n, const ull *inPtr, const uint32_t *inAdj, const uint32_t *outDeg, uint32_t *level,
                     uint32_t cur, uint8_t *labOut, ScanCtl *ctl, hipStream_t st);

// Streaming copy of `bytes` (multiple of 16) for counter calibration.
void launch_copy_probe(const void *src, void *dst, ull bytes, hipStream_t st);

// Fill helpers (kernels rather than memsets so they can sit in graphs).
void launch_fill_u8(uint8_t *p, uint8_t v, ull n, hipStream_t st);
void launch_fill_u32(uint32_t *p, uint32_t v, ull n, hipStream_t st);

}  // namespace subway

#endif  // SUBWAY_KERNELS_HPP
