// graph.hpp -- Graph<E> / GraphPR<E>: host CSR plus per-vertex state, the
// drop-in replacement of shared/graph.cuh:7-59.
//
// Public fields keep the reference's names and meaning (host arrays without
// prefix, device arrays with d_).  Differences a caller can see:
//  * offsets are 64-bit (nodePointer is ull[num_nodes+1], num_edges is ull),
//    so graphs with 2^32 or more edges load (the legacy u32 file header is
//    still read and written byte-exact);
//  * outDegree is correct for every vertex (the reference never writes
//    outDegree[num_nodes-2], shared/graph.cu:177-181);
//  * the pinned edge list is also mapped into the device address space
//    (d_edgeList) so kernels can gather from it without a staging copy;
//  * d_nodePointer mirrors nodePointer on the device;
//  * destructors release every allocation.
#ifndef SUBWAY_GRAPH_HPP
#define SUBWAY_GRAPH_HPP

#include <memory>
#include <mutex>
#include <string>

#include "globals.hpp"

namespace subway {

// Host CSR with pinned, device-mapped edges.  Shared between graph objects
// so that several runs (or a Graph and a GraphPR) reuse one loaded file.
template <class E>
struct HostCsr {
    ull n = 0;
    ull m = 0;
    ull *ptr = nullptr;     // n+1 offsets (ordinary host memory)
    E *edges = nullptr;     // m edges, hipHostMalloc'd (pinned + mapped)
    E *d_edges = nullptr;   // device address of `edges`
    uint *degree = nullptr; // n out-degrees
    uint maxDegree = 0;
    bool large = false;     // read from / destined for the 64-bit file variant
    bool pinned = false;    // edges live in pinned, device-mapped memory
    // Weighted graphs: every edge as one 32-bit word, destination |
    // (weight - packBase) << packShift (pinned, device-mapped), built on
    // first use when the vertex ids fit packShift bits and every weight minus
    // the smallest one (packBase, below 2^27) the remaining 32 - packShift.
    uint32_t *packed = nullptr;
    uint32_t *d_packed = nullptr;   // device address of `packed`
    uint32_t packShift = 0;         // 0: no packed copy (unweighted, or ids + weight span exceed 32 bits)
    uint32_t packBase = 0;          // smallest weight, added back when an edge is decoded
    uint32_t ensurePacked();        // builds the packed copy once; returns packShift
    // Weighted graphs that do not pack but whose weights fit a byte: the
    // destinations and the weights as two pinned, device-mapped arrays (5
    // bytes per edge instead of 8), built on first use.
    uint32_t *splitDst = nullptr, *d_splitDst = nullptr;
    uint8_t *splitW = nullptr, *d_splitW = nullptr;
    bool splitPinned = true;        // the split arrays are pinned and mapped (else ordinary memory, d_* null)
    bool ensureSplit();             // builds the split copy once; false when it does not apply
    // Moves pinned split arrays into ordinary memory (the process's pinned
    // memory is limited; a graph whose whole ranges move in the block layout
    // then gathers its compacted units on the host from these arrays).
    void demoteSplit();
    // After an unpinned read: split arrays, or the list moved into pinned memory
    void settleEdges(bool forceSplit);
    // Takes a pinned packed copy built elsewhere (a reader that never holds
    // the 8-byte list); ensurePacked() then returns `shift`.
    void adoptPacked(uint32_t *pinnedWords, uint32_t shift, uint32_t base);
    HostCsr() = default;
    HostCsr(const HostCsr &) = delete;
private:
    std::mutex packMu_;
    bool packTried_ = false;
    bool splitTried_ = false;
public:
    HostCsr &operator=(const HostCsr &) = delete;
    ~HostCsr();
    // allocates ptr, edges, degree; pinEdges = false keeps the edge list in
    // ordinary memory (a graph that moves only in the split format);
    // withEdges = false allocates no 8-byte list at all: the graph's edges
    // are then only the split arrays (allocateSplit), read through edgeAt()
    void allocate(ull n, ull m, bool pinEdges = true, bool withEdges = true);
    // edge i from whichever host layout the graph holds
    E edgeAt(ull i) const
    {
        if (edges) return edges[i];
        E e{};
        e.end = splitDst[i];
        if constexpr (sizeof(E) == 8) reinterpret_cast<uint32_t *>(&e)[1] = splitW[i];
        return e;
    }
    bool allocateSplit();          // pinned split arrays, filled by the caller; false if not allocated
    void finalize();               // degrees, maxDegree, device mapping
    // Vertex renumbering (reorder_by_in_degree): every array above then
    // describes the renumbered graph; perm[v] is the internal id of original
    // vertex v and iperm[p] the original id of internal vertex p (n entries
    // each, ordinary host memory).  Null while the graph keeps its own ids.
    // Runs take sources and return values in the original numbering.
    uint32_t *perm = nullptr;
    uint32_t *iperm = nullptr;
    // Block-compressed rows (the layout of subway_amd/csrc/blocks.hpp), built
    // on the device by build_blocks(): blk holds blkWords pinned words,
    // blkPtr[v] (n + 1 entries, ordinary memory) the word where row v starts,
    // blkCode the weight code (bits | base << 6).  rowsSorted: every row of
    // every host layout is in ascending destination order.
    uint32_t *blk = nullptr;
    ull *blkPtr = nullptr;
    ull blkWords = 0;
    uint32_t blkCode = 0;
    bool rowsSorted = false;
    // Split arrays in ordinary memory beside the block layout: the edges
    // [tailEdge, m) of the id suffix whose rows take nearly as many bytes in
    // the block layout as in the split format (low-degree rows) are also kept
    // pinned here, so whole ranges there move by DMA without staging.
    uint32_t *tailDst = nullptr;
    uint8_t *tailW = nullptr;
    ull tailEdge = 0;
};

// Renumbers the vertices in descending in-degree order (ties by ascending
// id), on the device, overwriting the graph's host arrays in place and
// setting perm / iperm.  Needs about twice the edge bytes of device memory.
// Call it before any run or session holds the graph.
template <class E>
void reorder_by_in_degree(HostCsr<E> &csr);

// Sorts every row by destination, on the device, rewriting each host layout
// the graph holds in place (results of every application are unchanged:
// only the order of a vertex's out-edges moves), then builds the
// block-compressed copy (blk, blkPtr, blkCode).  Needs about 1.5 times the
// graph's edge bytes of free device memory plus 13.5 GiB of sort scratch
// (less for graphs under 2^29 edges).  Call it before any run or session
// holds the graph.  unpinSplit: move pinned split arrays to ordinary memory
// before the block words are pinned (done anyway when pinning fails).
template <class E>
void build_blocks(HostCsr<E> &csr, bool unpinSplit = false);

// splitLayout: a .bwcsr graph that moves only in the split format (weights
// must fit a byte); weighted lists of 16 GiB or more choose it by themselves
// when they do not pack into one word per edge.
template <class E>
std::shared_ptr<HostCsr<E>> read_csr_file(const std::string &path, bool splitLayout = false);
template <class E>
void write_csr_file(const HostCsr<E> &csr, const std::string &path);

std::string file_extension(const std::string &fileName);

}  // namespace subway

template <class E>
class Graph {
public:
    std::string graphFilePath;
    bool isWeighted;
    bool isLarge;
    uint num_nodes;
    ull num_edges;
    ull *nodePointer;
    E *edgeList;
    uint *outDegree;
    bool *label1;
    bool *label2;
    uint *value;
    uint *d_outDegree;
    uint *d_value;
    bool *d_label1;
    bool *d_label2;
    ull *d_nodePointer;
    E *d_edgeList;          // device-mapped view of the pinned edgeList
    std::string graphFormat;
    std::shared_ptr<subway::HostCsr<E>> csr;

    Graph(std::string graphFilePath, bool isWeighted);
    // Wrap an already-built CSR (no file involved).
    explicit Graph(std::shared_ptr<subway::HostCsr<E>> csr);
    ~Graph();
    Graph(const Graph &) = delete;
    Graph &operator=(const Graph &) = delete;
    std::string GetFileExtension(std::string fileName);
    void AssignW8(uint w8, uint index);
    void ReadGraph();
    bool valueIsPinned() const { return valuePinned; }   // value is pinned host memory

private:
    bool valuePinned = false;   // value came from hipHostMalloc
    void attach();
};

template <class E>
class GraphPR {
public:
    std::string graphFilePath;
    bool isWeighted;
    bool isLarge;
    uint num_nodes;
    ull num_edges;
    ull *nodePointer;
    E *edgeList;
    uint *outDegree;
    float *value;
    float *delta;
    uint *d_outDegree;
    float *d_value;
    float *d_delta;    // the reference's FP32 residual (shared/graph.cuh:53); the engine keeps its own FP64
                       // residual while it runs (an FP32 sum at a hub rounds away shares below half its ulp)
                       // and leaves the final residual here, rounded to FP32
    ull *d_nodePointer;
    E *d_edgeList;
    std::string graphFormat;
    std::shared_ptr<subway::HostCsr<E>> csr;

    GraphPR(std::string graphFilePath, bool isWeighted);
    explicit GraphPR(std::shared_ptr<subway::HostCsr<E>> csr);
    ~GraphPR();
    GraphPR(const GraphPR &) = delete;
    GraphPR &operator=(const GraphPR &) = delete;
    std::string GetFileExtension(std::string fileName);
    void AssignW8(uint w8, uint index);
    void ReadGraph();
    bool valueIsPinned() const { return valuePinned; }   // value is pinned host memory

private:
    bool valuePinned = false;   // value came from hipHostMalloc
    void attach();
};

#endif  // SUBWAY_GRAPH_HPP
