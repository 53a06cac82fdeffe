// engine.cpp -- the plain-policy global-iteration loop shared by the ten
// applications (reference: subway/*-sync.cu and subway/*-async.cu).
//
// Initial state per application (host arrays, copied before the timed window):
//   BFS/SSSP  value INF, source 0            (bfs-sync.cu:28-36, sssp-async.cu:38-44)
//   SSWP      value 0, source INF            (sswp-sync.cu:28-36, sswp-async.cu:31-40)
//   CC        value v                        (cc-sync.cu:29-34, cc-async.cu:28-33)
//   PR        value 0, delta 0.15            (pr-sync.cu:33-40)
//   sync apps label label2[source] (CC: every label2); async apps label every
//   vertex for the first generate and then only the source (K8 in the
//   reference: sssp-async.cu:38-64); CC-async keeps every vertex labelled.
// Per partition: sync runs one frontier + advance round reading label2
// (moveUpLabels + kernel, bfs-sync.cu:69-71); async merges label2 into label1
// on the first round (mixLabels, sssp-async.cu:90) and alternates the label
// arrays round by round until a round selects nothing (sssp-async.cu:92-115).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <map>
#include <iostream>
#include <memory>
#include <vector>

#include "host_util.hpp"
#include "kernels.hpp"
#include "mlp.hpp"
#include "schedule.hpp"
#include "subway/engine.hpp"
#include "subway/gpu_error_check.hpp"
#include "subway/partitioner.hpp"
#include "subway/subgraph.hpp"
#include "subway/subgraph_generator.hpp"

namespace subway {

const char *algo_name(int algo)
{
    switch (algo) {
        case BFS: return "bfs";
        case CC: return "cc";
        case SSSP: return "sssp";
        case SSWP: return "sswp";
        case PR: return "pr";
        default: return "?";
    }
}

namespace {

typedef std::chrono::steady_clock clk;
inline double ms_since(clk::time_point t0)
{
    return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

// Start/stop event pairs recorded around launches when profiling; read out
// once at the end of a run.
class EventPairs {
public:
    ~EventPairs()
    {
        for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
    }
    void begin(hipStream_t s)
    {
        if (used_ + 2 > ev_.size()) {
            hipEvent_t a, b;
            SUBWAY_CHECK(hipEventCreate(&a));
            SUBWAY_CHECK(hipEventCreate(&b));
            ev_.push_back(a);
            ev_.push_back(b);
        }
        SUBWAY_CHECK(hipEventRecord(ev_[used_], s));
    }
    void end(hipStream_t s)
    {
        SUBWAY_CHECK(hipEventRecord(ev_[used_ + 1], s));
        used_ += 2;
    }
    void reset() { used_ = 0; }
    // total milliseconds of all pairs; call after the streams are drained
    double total(ull &count)
    {
        double ms = 0;
        for (size_t i = 0; i + 1 < used_; i += 2) {
            float t = 0;
            SUBWAY_CHECK(hipEventElapsedTime(&t, ev_[i], ev_[i + 1]));
            ms += t;
        }
        count = used_ / 2;
        return ms;
    }

private:
    std::vector<hipEvent_t> ev_;
    size_t used_ = 0;
};

// One thing to move into a device buffer: a whole partition, or one chunk of
// a vertex whose adjacency is larger than a buffer.
struct Unit {
    uint part;
    bool hub;
    ull chunk;          // hub chunk index
    bool dma;           // contiguous CSR slice: plain DMA
    ull hostEdge;       // first host edge index of a DMA transfer
    ull count;          // edges
    ull kbeg, kend;     // active-list range (gather)
    ull edgeBase;       // activeNodesPointer value at kbeg (gather)
    // the items the rounds run over
    ull n;              // number of items
    const uint *nodes;  // active-list slice, or null for a whole id range
    const ull *ptr;     // offsets of the items (n + 1 readable)
    ull fromEdge;       // subtracted from ptr to get buffer offsets
    uint idBase;        // first id of a whole range
    bool whole;         // whole id range resident (learned policy)
    // a DMA of every edge of the ids [idLo, idHi) (whole ranges, fully active
    // contiguous runs): the buffer then holds that id range's whole slice
    uint idLo, idHi;
    // ... moved in the block layout: words [blkLo, blkLo + blkWords) of the
    // graph's block copy (the rows of idLo .. idHi - 1)
    bool blocked;
    ull blkLo, blkWords;
};

// One compute, one copy and one read-back stream per device, shared by
// every engine of the process.  HIP maps streams onto a few hardware queues
// (four by default); with streams per session, a second session would put
// its copy and compute streams on one queue and lose the transfer/compute
// overlap.  Runs of different sessions on one device therefore serialise on
// the streams.  The read-back stream carries the early copy of a run's
// results to the host (device to host, beside the copy stream's host to
// device transfers).
//
// The pairs are destroyed explicitly, while the HIP runtime is still whole:
// by release_shared_streams() (the C ABI's subway_shutdown, which the Python
// package calls from its exit hook after closing every session and graph),
// or else by an atexit handler registered when the first pair is made.  That
// handler runs before the runtime's own teardown (atexit order is the
// reverse of registration, and the runtime was initialised before the first
// stream existed), so the runtime never has to reclaim live streams of this
// library during process exit.
static std::mutex g_streamMu;
struct DeviceStreams {
    hipStream_t compute, copy, back;
};
static std::map<int, DeviceStreams> *g_streamPairs = nullptr;

static void destroy_stream_pairs()
{
    std::lock_guard<std::mutex> lock(g_streamMu);
    if (!g_streamPairs) return;
    for (auto &kv : *g_streamPairs) {
        int cur = 0;
        if (hipGetDevice(&cur) != hipSuccess) cur = kv.first;
        (void)hipSetDevice(kv.first);
        for (hipStream_t s : {kv.second.compute, kv.second.copy, kv.second.back}) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        (void)hipSetDevice(cur);
    }
    delete g_streamPairs;
    g_streamPairs = nullptr;
}

static void shared_streams(hipStream_t &compute, hipStream_t &copy, hipStream_t &back)
{
    static bool atexitRegistered = false;
    int dev = 0;
    SUBWAY_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_streamMu);
    if (!g_streamPairs) g_streamPairs = new std::map<int, DeviceStreams>();
    auto it = g_streamPairs->find(dev);
    if (it == g_streamPairs->end()) {
        DeviceStreams s{};
        SUBWAY_CHECK(hipStreamCreateWithFlags(&s.compute, hipStreamNonBlocking));
        SUBWAY_CHECK(hipStreamCreateWithFlags(&s.copy, hipStreamNonBlocking));
        SUBWAY_CHECK(hipStreamCreateWithFlags(&s.back, hipStreamNonBlocking));
        it = g_streamPairs->emplace(dev, s).first;
        if (!atexitRegistered) {
            std::atexit(destroy_stream_pairs);
            atexitRegistered = true;
        }
    }
    compute = it->second.compute;
    copy = it->second.copy;
    back = it->second.back;
}

template <class E, class G>
class Engine {
public:
    // Non-owning: runs on a graph object the caller keeps (drop-in apps).
    Engine(G &g, const RunOptions &o) : g_(&g), o_(o) { setup(); }
    // Owning: builds the graph's device state from a shared CSR (sessions).
    Engine(std::shared_ptr<HostCsr<E>> csr, const RunOptions &o) : own_(new G(std::move(csr))), o_(o)
    {
        g_ = own_.get();
        setup();
    }
    G &graph() { return *g_; }

    void setup()
    {
        if (o_.numBuffers != 1 && o_.numBuffers != 2) throw Error(ERR_ARG, "numBuffers must be 1 or 2");
        if (o_.budgetBytes)
            sub_.reset(new Subgraph<E>(g_->num_nodes, g_->num_edges, o_.budgetBytes, o_.numBuffers));
        else
            sub_.reset(new Subgraph<E>(g_->num_nodes, g_->num_edges));
        gen_.reset(new SubgraphGenerator<E>(*g_));
        part_.reset(new Partitioner<E>());
        shared_streams(sc_, sx_, sd_);
        gen_->stream = sc_;
        part_->stream = sc_;
        for (int b = 0; b < 2; b++) {
            SUBWAY_CHECK(hipEventCreateWithFlags(&evLoaded_[b], hipEventDisableTiming));
            SUBWAY_CHECK(hipEventCreateWithFlags(&evFree_[b], hipEventDisableTiming));
            SUBWAY_CHECK(hipEventCreateWithFlags(&evStaged_[b], hipEventDisableTiming));
        }
        SUBWAY_CHECK(hipEventCreate(&evX0_));
        SUBWAY_CHECK(hipEventCreate(&evX1_));
        maxPart_ = sub_->max_partition_size;
        if constexpr (std::is_same<E, OutEdgeWeighted>::value) {
            // weighted edges travel and sit in the buffers as one word each
            // when ids and weights fit: twice the edges per buffer, half the
            // bytes per transferred edge
            if (!o_.unpackedEdges && g_->csr->ensurePacked() != 0) {
                packCode_ = g_->csr->packShift | (g_->csr->packBase << 5);
                fmt_ = EDGE_PACKED;
                eb_ = sizeof(uint32_t);
                maxPart_ = std::min<ull>(maxPart_ * (sizeof(E) / sizeof(uint32_t)), 0x7FFFF800ull);
            } else if (!o_.unpackedEdges && g_->csr->ensureSplit()) {
                // ids too wide to pack, weights in a byte: a u32 destination
                // array and a u8 weight array per buffer, 5 bytes per edge
                ull cap = maxPart_ * sizeof(E) / 5;
                cap = std::min<ull>(cap - cap % ADV_CHUNK, 0x7FFFF800ull);
                if (cap > maxPart_) {
                    fmt_ = EDGE_SPLIT;
                    eb_ = 5;
                    maxPart_ = cap;
                    packCode_ = (uint32_t)cap;
                }
            }
        }
        bufBytes_ = (ull)sub_->max_partition_size * sizeof(E);
        edgesMapped_ = fmt_ == EDGE_SPLIT ? g_->csr->d_splitDst != nullptr : true;
        if (g_->csr->blk && !o_.noBlocks && !o_.batch && g_->num_edges > 0) {
            // whole id ranges move in the block layout; the frontier finds a
            // row's words through the graph's row offsets
            blkOk_ = true;
            blkCode_ = g_->csr->blkCode;
            const size_t bytes = ((size_t)g_->num_nodes + 1) * sizeof(ull);
            SUBWAY_CHECK(hipMalloc(&d_blk_, bytes));
            SUBWAY_CHECK(hipMemcpy(d_blk_, g_->csr->blkPtr, bytes, hipMemcpyHostToDevice));
            SUBWAY_CHECK(hipDeviceSynchronize());   // a pageable copy may return before its DMA ends
            void *dev = nullptr;
            if (hipHostGetDevicePointer(&dev, g_->csr->blk, 0) == hipSuccess && dev) {
                blkDev_ = (const uint32_t *)dev;
                SUBWAY_CHECK(hipMalloc(&d_wpre_, bytes));
                alloc_scan_scratch(rw_, ((ull)g_->num_nodes + GEN_TILE - 1) / GEN_TILE);
            } else {
                (void)hipGetLastError();   // not mapped: gathered units move their plain edges
            }
        }
        // a graph that fits one buffer moves once per run and stays there
        resident_ = !o_.noResident && !o_.batch && g_->num_edges > 0 && g_->num_edges <= maxPart_;
        if constexpr (!isPR) {
            if (resident_ && o_.algo == BFS && !o_.pushOnly) buildInEdges();
        }
        if (!g_->csr->pinned && fmt_ != EDGE_SPLIT && fmt_ != EDGE_PACKED)
            throw Error(ERR_ARG, "this graph keeps no pinned 8-byte edge list: it moves only in the split format "
                                 "or as packed words (unpacked_edges must be 0)");
        alloc_scan_scratch(fr_, ((ull)g_->num_nodes + FR_TILE - 1) / FR_TILE + 1);
        ull maxEntries = std::min<ull>((ull)g_->num_nodes, maxPart_) + 1;
        SUBWAY_CHECK(hipMalloc(&entries_, maxEntries * sizeof(Entry)));
        ull maxWork = std::max<ull>(maxPart_, (ull)g_->csr->maxDegree);
        const ull chunk = o_.batch ? (ull)BATCH_CHUNK_EDGES : (ull)ADV_CHUNK;
        SUBWAY_CHECK(hipMalloc(&chunkMap_, (maxWork / chunk + 2) * sizeof(uint32_t)));
        SUBWAY_CHECK(hipHostMalloc((void **)&h_ctl_, sizeof(ScanCtl), 0));
        SUBWAY_CHECK(hipHostMalloc((void **)&h_err_, 4 * sizeof(uint32_t), 0));
        std::memset(h_err_, 0, 4 * sizeof(uint32_t));
        fr_.spinLimit = o_.spinLimit ? o_.spinLimit : dev::SPIN_LIMIT_DEFAULT;
        if constexpr (isPR) {
            SUBWAY_CHECK(hipMalloc(&d_sum_, (size_t)g_->num_nodes * sizeof(double)));
            SUBWAY_CHECK(hipMalloc(&d_resid_, (size_t)g_->num_nodes * sizeof(double)));
        }
        if (g_->csr->perm) {
            const ull n = g_->num_nodes;
            SUBWAY_CHECK(hipMalloc(&d_perm_, n * sizeof(uint32_t)));
            SUBWAY_CHECK(hipMemcpy(d_perm_, g_->csr->perm, n * sizeof(uint32_t), hipMemcpyHostToDevice));
            if (o_.algo == CC) {
                SUBWAY_CHECK(hipMalloc(&d_iperm_, n * sizeof(uint32_t)));
                SUBWAY_CHECK(hipMemcpy(d_iperm_, g_->csr->iperm, n * sizeof(uint32_t), hipMemcpyHostToDevice));
            }
            // batched rows (256 bytes a vertex) go out in chunks of 2^20 vertices
            unpermRows_ = o_.batch ? std::min<ull>(n, 1ull << 20) : n;
            SUBWAY_CHECK(hipMalloc(&d_unperm_, unpermRows_ * (o_.batch ? BATCH_LANES : 1) * sizeof(uint32_t)));
        }
        gen_->scanScratch().spinLimit = fr_.spinLimit;
        if (!o_.batch) setupEarlyReadback();
        int dev = 0;
        SUBWAY_CHECK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        SUBWAY_CHECK(hipGetDeviceProperties(&prop, dev));
        advGrid_ = o_.advanceGrid > 0 ? o_.advanceGrid : prop.multiProcessorCount * 4;
        if (o_.policy == POLICY_LEARNED) setupRanges();
        if (o_.batch) {
            setupBatch(prop, maxEntries);
            return;
        }
        if (o_.persistent) {
            int mx = rounds_max_grid(o_.algo, fmt_);
            if (blkOk_) mx = std::min(mx, rounds_max_grid(o_.algo, EDGE_BLOCKS));
            roundsGrid_ = std::min(mx, o_.roundsGrid > 0 ? o_.roundsGrid : prop.multiProcessorCount * 4);
            if (roundsGrid_ > 0) {
                SUBWAY_CHECK(hipMalloc(&pc_, sizeof(dev::PersistCtl)));
                SUBWAY_CHECK(hipMemset(pc_, 0, sizeof(dev::PersistCtl)));
                SUBWAY_CHECK(hipDeviceSynchronize());
            }
        }
    }

    // Batched k-source state: dist rows [V][64], two 64-bit mask arrays,
    // per-entry masks, and a persistent grid for the batched rounds.
    void setupBatch(const hipDeviceProp_t &prop, ull maxEntries)
    {
        if constexpr (isPR) {
            throw Error(ERR_ARG, "batched runs support BFS, SSSP and SSWP");
        } else {
            if (o_.algo != BFS && o_.algo != SSSP && o_.algo != SSWP)
                throw Error(ERR_ARG, "batched runs support BFS, SSSP and SSWP");
            if (o_.mode != ASYNC) throw Error(ERR_ARG, "batched runs are async");
            const ull n = g_->num_nodes;
            SUBWAY_CHECK(hipMalloc(&d_dist_, n * BATCH_LANES * sizeof(uint32_t)));
            SUBWAY_CHECK(hipMalloc(&d_m1_, n * sizeof(ull)));
            SUBWAY_CHECK(hipMalloc(&d_m2_, n * sizeof(ull)));
            SUBWAY_CHECK(hipMalloc(&d_entryMask_, maxEntries * sizeof(ull)));
            SUBWAY_CHECK(hipMalloc(&d_sources_, BATCH_LANES * sizeof(uint32_t)));
            SUBWAY_CHECK(hipHostMalloc((void **)&h_sources_, BATCH_LANES * sizeof(uint32_t), 0));
            setBatchUnroll(o_.batchUnroll > 0 ? o_.batchUnroll : 8);
            SUBWAY_CHECK(hipMalloc(&pc_, sizeof(dev::PersistCtl)));
            SUBWAY_CHECK(hipMemset(pc_, 0, sizeof(dev::PersistCtl)));
            SUBWAY_CHECK(hipDeviceSynchronize());
            (void)prop;
        }
    }

public:
    void setBatchUnroll(int u)
    {
        int dev = 0;
        SUBWAY_CHECK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        SUBWAY_CHECK(hipGetDeviceProperties(&prop, dev));
        const int mx = batch_max_grid(o_.algo, fmt_, u);
        if (mx <= 0) throw Error(ERR_ARG, "this device cannot run the persistent batched rounds kernel");
        batchUnroll_ = u;
        batchGrid_ = std::min(mx, o_.roundsGrid > 0 ? o_.roundsGrid : prop.multiProcessorCount * 4);
    }
    int batchUnroll() const { return batchUnroll_; }
    RunStats runBatch(const uint *sources, int k);
    // scheduler network of the learned policy (MLP_PARAMS floats, host); null removes it
    void setPolicy(const float *params)
    {
        if (!params) {
            if (d_params_) SUBWAY_CHECK(hipFree(d_params_));
            d_params_ = nullptr;
            return;
        }
        if (!d_bounds_) throw Error(ERR_ARG, "a scheduler network needs the learned policy");
        if (!d_params_) SUBWAY_CHECK(hipMalloc(&d_params_, MLP_PARAMS * sizeof(float)));
        SUBWAY_CHECK(hipMemcpy(d_params_, params, MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice));
    }
    // rows of the last batched run: n x 64 u32, vertex-major
    const uint32_t *batchRows() const { return h_dist_; }

    ~Engine()
    {
        if (sc_) (void)hipStreamSynchronize(sc_);
        if (sx_) (void)hipStreamSynchronize(sx_);
        if (d_bounds_) (void)hipFree(d_bounds_);
        if (d_pol_) (void)hipFree(d_pol_);
        if (h_pol_) (void)hipHostFree(h_pol_);
        if (d_pstate_) (void)hipFree(d_pstate_);
        if (d_params_) (void)hipFree(d_params_);
        if (d_dist_) (void)hipFree(d_dist_);
        if (d_m1_) (void)hipFree(d_m1_);
        if (d_m2_) (void)hipFree(d_m2_);
        if (d_entryMask_) (void)hipFree(d_entryMask_);
        if (d_sources_) (void)hipFree(d_sources_);
        if (h_sources_) (void)hipHostFree(h_sources_);
        if (h_dist_) (void)hipHostFree(h_dist_);
        free_scan_scratch(fr_);
        if (entries_) (void)hipFree(entries_);
        if (chunkMap_) (void)hipFree(chunkMap_);
        if (h_ctl_) (void)hipHostFree(h_ctl_);
        if (h_err_) (void)hipHostFree(h_err_);
        if (d_sum_) (void)hipFree(d_sum_);
        if (d_resid_) (void)hipFree(d_resid_);
        if (d_perm_) (void)hipFree(d_perm_);
        if (d_iperm_) (void)hipFree(d_iperm_);
        if (d_unperm_) (void)hipFree(d_unperm_);
        if (d_inPtr_) (void)hipFree(d_inPtr_);
        if (d_inAdj_) (void)hipFree(d_inAdj_);
        if (d_blk_) (void)hipFree(d_blk_);
        if (d_wpre_) {
            (void)hipFree(d_wpre_);
            free_scan_scratch(rw_);
        }
        if (h_wb_) (void)hipHostFree(h_wb_);
        if (iterExec_) (void)hipGraphExecDestroy(iterExec_);
        if (pc_) (void)hipFree(pc_);
        for (int b = 0; b < 2; b++) {
            (void)hipEventDestroy(evLoaded_[b]);
            (void)hipEventDestroy(evFree_[b]);
            (void)hipEventDestroy(evStaged_[b]);
            if (stage_[b]) (void)hipHostFree(stage_[b]);
        }
        if (h_ids_) (void)hipHostFree(h_ids_);
        if (sd_) (void)hipStreamSynchronize(sd_);
        if (d_snap_) (void)hipFree(d_snap_);
        if (d_snapInt_) (void)hipFree(d_snapInt_);
        if (d_chg_) (void)hipFree(d_chg_);
        if (d_chgCount_) (void)hipFree(d_chgCount_);
        if (h_chg_) (void)hipHostFree(h_chg_);
        if (h_chgCount_) (void)hipHostFree(h_chgCount_);
        if (evSnap_) (void)hipEventDestroy(evSnap_);
        if (evBack_) (void)hipEventDestroy(evBack_);
        (void)hipEventDestroy(evX0_);
        (void)hipEventDestroy(evX1_);
        part_.reset();
        gen_.reset();
        sub_.reset();
        // the streams are the process's shared pair (see shared_streams)
    }

    RunStats run(uint source);

private:
    std::unique_ptr<G> own_;
    G *g_ = nullptr;
    RunOptions o_;
    uint src_ = 0;
    std::unique_ptr<Subgraph<E>> sub_;
    std::unique_ptr<SubgraphGenerator<E>> gen_;
    std::unique_ptr<Partitioner<E>> part_;
    hipStream_t sc_ = nullptr, sx_ = nullptr, sd_ = nullptr;
    hipEvent_t evLoaded_[2], evFree_[2], evX0_, evX1_;
    ScanScratch fr_;
    Entry *entries_ = nullptr;
    uint32_t *chunkMap_ = nullptr;
    ScanCtl *h_ctl_ = nullptr;
    uint32_t *h_err_ = nullptr;   // pinned: error words of the frontier, generator and rounds control blocks
    // renumbered graph: perm (original -> internal), iperm (CC's initial
    // values), and the result in the original order before its copy out
    uint32_t *d_perm_ = nullptr, *d_iperm_ = nullptr, *d_unperm_ = nullptr;
    ull unpermRows_ = 0;          // vertices d_unperm_ holds (batched rows are chunked)
    void copyValuesOut(const void *dValues, void *hValues, uint32_t width);
    double *d_sum_ = nullptr;     // PR: FP64 running sum of each vertex's absorbed residual
    double *d_resid_ = nullptr;   // PR: FP64 residual (delta) while a run is under way
    // BFS on a resident graph: the in-edges (CSC) for bottom-up steps
    ull *d_inPtr_ = nullptr;
    uint32_t *d_inAdj_ = nullptr;
    void buildInEdges();
    void bfsDirectionOptimizing(const Unit &un, int buf);
    bool pullCounted_ = false;
    bool broken_ = false;         // a device wait gave up: the device state is no longer trusted
    ull maxPart_ = 0;       // edges per device buffer (in the buffer's edge format)
    int fmt_ = (int)sizeof(E);          // EdgeFormat of the device buffers and transfers
    int eb_ = (int)sizeof(E);           // bytes per transferred / buffered edge
    uint32_t packCode_ = 0;            // EDGE_PACKED: shift | weight base << 5
    // block layout: row word offsets on the device, the weight code, and the
    // bytes of a device buffer; bufBlocked_ = the buffer holds a block slice
    ull *d_blk_ = nullptr;
    bool blkOk_ = false;
    uint32_t blkCode_ = 0;
    // gathered block units: the GPU reads the rows of a unit's active
    // vertices from the graph's mapped block words (blkDev_) into the buffer
    // in active-list order; d_wpre_[k] = words of the rows of active items
    // before k (a scan per global iteration, when a unit needs it)
    const uint32_t *blkDev_ = nullptr;
    ull *d_wpre_ = nullptr;
    ScanScratch rw_{};
    bool wpreValid_ = false;
    ull *h_wb_ = nullptr;        // pinned: the scan's values at the units' item bounds
    size_t wbCap_ = 0;
    static constexpr ull kGatherBlockedMinBytes = 16ull << 20;
    void markGatheredBlocked();
    ull bufBytes_ = 0;
    bool bufBlocked_[2] = {false, false};
    // the GPU can read the host edges of the transfer format (mapped pinned
    // memory); otherwise compacted units are gathered by host threads
    bool edgesMapped_ = true;
    bool blkDiag_ = false;
    void markBlocked(Unit &u) const;
    int fmtOf(const Unit &u) const { return u.blocked ? (int)EDGE_BLOCKS : fmt_; }
    uint32_t codeOf(const Unit &u) const { return u.blocked ? blkCode_ : packCode_; }
    int advGrid_ = 0;
    RunStats st_;
    std::vector<Unit> units_;
    EventPairs pAdv_, pFr_, pXfer_, pGen_, pRounds_, pRoundsBlk_;   // pRoundsBlk_: rounds launches on block units
    dev::PersistCtl *pc_ = nullptr;
    int roundsGrid_ = 0;    // 0: persistent rounds kernel unavailable
    // batched state
    int batchK_ = 0, batchUnroll_ = 8, batchGrid_ = 0;
    uint32_t *d_dist_ = nullptr, *d_sources_ = nullptr, *h_sources_ = nullptr, *h_dist_ = nullptr;
    ull *d_m1_ = nullptr, *d_m2_ = nullptr, *d_entryMask_ = nullptr;

    static constexpr bool isPR = std::is_same<G, GraphPR<E>>::value;

    void initState();
    void generateAndPartition();
    void errorCopies();
    hipGraphExec_t iterExec_ = nullptr;   // batched runs: the recorded start of an iteration
    bool resident_ = false;               // the whole graph fits buffer 0
    bool residentLoaded_ = false;         // ... and has been moved there in this run
    // host edge slice [hostEdge, hostEdge + count) each buffer holds from a
    // DMA of this run (count 0: unknown); host edges never change, so a unit
    // asking for the same slice in the same buffer needs no copy
    ull bufHostEdge_[2] = {0, 0}, bufCount_[2] = {0, 0};
    // id range whose whole edge slice the buffer holds (idHi > idLo), and the
    // rotation of the unit -> buffer map of the current iteration
    uint bufIdLo_[2] = {0, 0}, bufIdHi_[2] = {0, 0};
    size_t bufShift_ = 0;
    void buildUnits();
    void issue(size_t u);
    void compute(size_t u);
    void frontier(int first, uint8_t *labA, uint8_t *labB, const Unit &un);
    void advance(uint8_t *labOut, ull w0, ull w1, ull shift, int buf, const Unit &un);
    void asyncRounds(const Unit &un, int buf);
    // gathers the unit's active adjacency, elements of esz bytes from src, into
    // the device array dstDev, through the two pinned staging slots
    void hostGather(const Unit &u, const char *src, size_t esz, char *dstDev);
    // copies a contiguous host range in ordinary memory to the device through the slots
    void stagedCopy(const char *host, char *dstDev, size_t bytes);
    char *stageSlot(int &s);
    static constexpr size_t kStageBytes = 64ull << 20;
    char *stage_[2] = {nullptr, nullptr};
    int stageSlot_ = 0;
    std::vector<ull> gatherPre_;
    // host edge array in the transfer format (pinned)
    const char *hostEdges() const
    {
        return fmt_ == EDGE_PACKED ? (const char *)g_->csr->packed : (const char *)g_->edgeList;
    }
    // items of a compacted unit: a slice of the active list
    void setItems(Unit &u, ull fromNode, ull cnt, ull fromEdge)
    {
        u.n = cnt;
        u.nodes = sub_->d_activeNodes + fromNode;
        u.ptr = sub_->d_activeNodesPointer + fromNode;
        u.fromEdge = fromEdge;
        u.idBase = 0;
        u.whole = false;
    }
    int roundsCap_ = 1 << 16;
    // learned policy: static id ranges of at most one buffer of edges
    std::vector<uint> bounds_;
    uint32_t *d_bounds_ = nullptr;
    char *d_pol_ = nullptr, *h_pol_ = nullptr;   // PolicyOut, then P x 4 range stats
    ull *h_rstats_ = nullptr;                     // range stats inside h_pol_
    size_t polBytes_ = 0;
    PolicyState *d_pstate_ = nullptr;
    float *d_params_ = nullptr;                   // scheduler network; null: fixed theta / cap
    bool onPlain_ = false;                        // learned policy fell back to the plain schedule
    float predEdges_ = -1.f;                      // last prediction of the next active-edge fraction
    uint judged_ = 0;
    double theta_ = 2.0;    // density at or above which a range moves whole
    void applyPolicy(IterRecord &rec);
    void setupRanges();
    void buildUnitsLearned();
    void pushHubUnits(uint part, uint v, ull fromNode, ull fromEdge, ull edges);
    void computeBatch(const Unit &u, int b);
    RunStats loop(const std::chrono::steady_clock::time_point &t0);
    void resetRun();
    void checkWaits();
    // statistics counters at the start of the run (host mirror of the scan control block)
    ScanCtl ctl0_{};
    void startStats();
    void finishStats();
    uint *pinnedIds(ull n)
    {
        if (n > idsCap_) {
            if (h_ids_) SUBWAY_CHECK(hipHostFree(h_ids_));
            idsCap_ = std::max<ull>(n, (ull)g_->num_nodes);
            SUBWAY_CHECK(hipHostMalloc((void **)&h_ids_, idsCap_ * sizeof(uint), 0));
        }
        return h_ids_;
    }
    uint *h_ids_ = nullptr;
    ull idsCap_ = 0;
    bool hostIdsValid_ = false;
    hipEvent_t evStaged_[2];
    ull deviceRounds();
    // Early read-back of single-source results: once the active edges have
    // fallen to 1/64 of their peak (and an iteration follows), the values (original order)
    // are copied to d_snap_ and from there to the host result array on the
    // read-back stream, while the later iterations run; at the end only the
    // values that changed since then are copied (a list of at most chgCap_
    // (vertex, value) pairs; more than that: the whole array again).
    bool earlyOk_ = false;      // buffers allocated: the session may read back early
    bool earlyRun_ = false;     // this run is a single-source run
    bool early_ = false;        // this run's early copy was made
    uint32_t *d_snap_ = nullptr, *d_snapInt_ = nullptr, *d_chgCount_ = nullptr, *h_chgCount_ = nullptr;
    uint2 *d_chg_ = nullptr, *h_chg_ = nullptr;
    uint32_t chgCap_ = 0;
    uint32_t *valueDev_ = nullptr;   // device address of the (mapped, pinned) host result array
    hipEvent_t evSnap_ = nullptr, evBack_ = nullptr;
    static constexpr ull kEarlyMinVertices = 1ull << 26;
    static constexpr int kStoreGrid = 64;   // workgroups of the early copy's store kernel
    void setupEarlyReadback();
    void startEarlyReadback();
    void finishValues();
};

template <class E, class G>
void Engine<E, G>::setupEarlyReadback()
{
    const ull n = g_->num_nodes;
    if constexpr (isPR) return;   // PR's values exist only after the last iteration
    // small graphs copy their results in well under a millisecond; a result
    // array in ordinary memory would make the early copy synchronous
    if (o_.earlyReadback < 0 || n == 0 || !g_->valueIsPinned()) return;
    if (o_.earlyReadback == 0 && n < kEarlyMinVertices) return;
    chgCap_ = (uint32_t)std::min<ull>(n, o_.earlyReadback > 0 ? (ull)o_.earlyReadback : 1ull << 22);
    if (hipHostMalloc((void **)&h_chg_, (size_t)chgCap_ * sizeof(uint2), 0) != hipSuccess ||
        hipHostMalloc((void **)&h_chgCount_, sizeof(uint32_t), 0) != hipSuccess) {
        (void)hipGetLastError();   // pinned memory is short: results are copied at the end
        if (h_chg_) (void)hipHostFree(h_chg_);
        h_chg_ = nullptr;
        h_chgCount_ = nullptr;
        return;
    }
    if (hipHostGetDevicePointer((void **)&valueDev_, g_->value, 0) != hipSuccess || !valueDev_ ||
        ((uintptr_t)valueDev_ & 15u)) {
        (void)hipGetLastError();   // the result array is not mapped: results are copied at the end
        (void)hipHostFree(h_chg_);
        (void)hipHostFree(h_chgCount_);
        h_chg_ = nullptr;
        h_chgCount_ = nullptr;
        valueDev_ = nullptr;
        return;
    }
    SUBWAY_CHECK(hipMalloc(&d_snap_, n * sizeof(uint32_t)));
    if (d_perm_) {
        // renumbered: the changes are found in the internal order (a second
        // copy of the values) and named through the inverse map
        SUBWAY_CHECK(hipMalloc(&d_snapInt_, n * sizeof(uint32_t)));
        if (!d_iperm_) {
            SUBWAY_CHECK(hipMalloc(&d_iperm_, n * sizeof(uint32_t)));
            SUBWAY_CHECK(hipMemcpy(d_iperm_, g_->csr->iperm, n * sizeof(uint32_t), hipMemcpyHostToDevice));
            SUBWAY_CHECK(hipDeviceSynchronize());   // a pageable copy may return before its DMA ends
        }
    }
    SUBWAY_CHECK(hipMalloc(&d_chg_, (size_t)chgCap_ * sizeof(uint2)));
    SUBWAY_CHECK(hipMalloc(&d_chgCount_, sizeof(uint32_t)));
    SUBWAY_CHECK(hipEventCreateWithFlags(&evSnap_, hipEventDisableTiming));
    SUBWAY_CHECK(hipEventCreateWithFlags(&evBack_, hipEventDisableTiming));
    earlyOk_ = true;
}

// The values as they are now (the later iterations change few of them), in
// the original order, to d_snap_ (compute stream), then to the host result
// array by a store kernel on the read-back stream (not a copy engine: the
// host's small read-backs of the later iterations must not queue behind
// it).  Nothing writes d_snap_ until the next run.
template <class E, class G>
void Engine<E, G>::startEarlyReadback()
{
    const ull n = g_->num_nodes;
    if (d_perm_) {
        // the internal-order copy the end compares against, and the values
        // in the original order for the store (a gather through perm inside
        // the store, by its 64 workgroups, took 3 ms longer at scale 28 than
        // this full-device pass)
        SUBWAY_CHECK(hipMemcpyAsync(d_snapInt_, g_->d_value, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, sc_));
        launch_unpermute((const uint32_t *)g_->d_value, d_perm_, 0, n, 1, d_snap_, sc_);
    } else {
        SUBWAY_CHECK(hipMemcpyAsync(d_snap_, g_->d_value, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, sc_));
    }
    SUBWAY_CHECK(hipEventRecord(evSnap_, sc_));
    SUBWAY_CHECK(hipStreamWaitEvent(sd_, evSnap_, 0));
    launch_store_host(d_snap_, valueDev_, n, kStoreGrid, sd_);
    SUBWAY_CHECK(hipEventRecord(evBack_, sd_));
    early_ = true;
}

// The run's values into the host result array: all of them, or after an
// early copy the ones that changed since (patched on the host once the early
// copy has landed).
template <class E, class G>
void Engine<E, G>::finishValues()
{
    if (!early_) {
        copyValuesOut(g_->d_value, g_->value, 1);
        return;
    }
    early_ = false;
    const ull n = g_->num_nodes;
    launch_fill_u32(d_chgCount_, 0u, 1, sc_);
    launch_value_changes((const uint32_t *)g_->d_value, d_perm_ ? d_iperm_ : nullptr, d_perm_ ? d_snapInt_ : d_snap_, n,
                         d_chgCount_, d_chg_, chgCap_, sc_);
    SUBWAY_CHECK(hipMemcpyAsync(h_chgCount_, d_chgCount_, sizeof(uint32_t), hipMemcpyDeviceToHost, sc_));
    SUBWAY_CHECK(hipStreamSynchronize(sc_));
    const uint32_t c = *h_chgCount_;
    SUBWAY_CHECK(hipEventSynchronize(evBack_));   // the early copy has landed
    st_.earlyReadback = 1;
    st_.valuesPatched = c;
    if (c > chgCap_) {
        st_.earlyReadback = 2;
        copyValuesOut(g_->d_value, g_->value, 1);
        return;
    }
    if (c == 0) return;
    SUBWAY_CHECK(hipMemcpyAsync(h_chg_, d_chg_, (size_t)c * sizeof(uint2), hipMemcpyDeviceToHost, sc_));
    SUBWAY_CHECK(hipStreamSynchronize(sc_));
    uint32_t *val = (uint32_t *)g_->value;
    const uint2 *chg = h_chg_;
    if (c < (1u << 16)) {
        for (uint32_t i = 0; i < c; i++) val[chg[i].x] = chg[i].y;
    } else {
        parallel_for(c, [&](ull a, ull b, int) {
            for (ull i = a; i < b; i++) val[chg[i].x] = chg[i].y;
        });
    }
}

template <class E, class G>
void Engine<E, G>::initState()
{
    const uint n = g_->num_nodes;
    if constexpr (isPR) {
        launch_init_pr_state(n, d_sum_, d_resid_, o_.initPR, sc_);
    } else {
        launch_init_label_state(o_.algo, o_.mode == ASYNC, n, g_->d_value, (uint8_t *)g_->d_label1,
                                (uint8_t *)g_->d_label2, src_, d_iperm_, sc_);
    }
}

template <class E, class G>
void Engine<E, G>::errorCopies()
{
    SUBWAY_CHECK(hipMemcpyAsync(&h_err_[0], &fr_.ctl->error, sizeof(uint32_t), hipMemcpyDeviceToHost, sc_));
    SUBWAY_CHECK(hipMemcpyAsync(&h_err_[1], &gen_->scanScratch().ctl->error, sizeof(uint32_t),
                                hipMemcpyDeviceToHost, sc_));
    if (pc_) SUBWAY_CHECK(hipMemcpyAsync(&h_err_[2], &pc_->error, sizeof(uint32_t), hipMemcpyDeviceToHost, sc_));
}

template <class E, class G>
void Engine<E, G>::generateAndPartition()
{
    hostIdsValid_ = false;
    wpreValid_ = false;
    if (batchK_ && !o_.profile && !d_bounds_) {
        // A batched iteration starts with the same launches and copies every
        // time (union generate, error words, partition table): recorded once
        // as a HIP graph (no memset nodes) and replayed.  Not with the
        // learned policy (its kernel takes the iteration index) or when
        // profiling (events around the launches).
        if (!iterExec_) {
            hipGraph_t gr = nullptr;
            SUBWAY_CHECK(hipStreamBeginCapture(sc_, hipStreamCaptureModeThreadLocal));
            try {
                launch_generate_masks(g_->num_nodes, d_m1_, d_m2_, g_->d_outDegree, sub_->d_activeNodes,
                                      sub_->d_activeNodesPointer, gen_->scanScratch(), sc_);
                errorCopies();
                part_->enqueue(*sub_, gen_->deviceTotals(), maxPart_);
            } catch (...) {
                (void)hipStreamEndCapture(sc_, &gr);
                if (gr) (void)hipGraphDestroy(gr);
                throw;
            }
            SUBWAY_CHECK(hipStreamEndCapture(sc_, &gr));
            const hipError_t e = hipGraphInstantiate(&iterExec_, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            SUBWAY_CHECK(e);
        }
        SUBWAY_CHECK(hipGraphLaunch(iterExec_, sc_));
        part_->collect(*sub_, gen_->deviceTotals(), maxPart_);
        checkWaits();
        sub_->numActiveNodes = (uint)part_->numActiveNodes;
        sub_->numActiveEdges = part_->numActiveEdges;
        return;
    }
    if (o_.profile) pGen_.begin(sc_);
    if (batchK_)
        launch_generate_masks(g_->num_nodes, d_m1_, d_m2_, g_->d_outDegree, sub_->d_activeNodes,
                              sub_->d_activeNodesPointer, gen_->scanScratch(), sc_);
    else if constexpr (isPR)
        launch_generate_pr(g_->num_nodes, d_resid_, (double)o_.acc, g_->d_outDegree, sub_->d_activeNodes,
                           sub_->d_activeNodesPointer, gen_->scanScratch(), sc_);
    else
        gen_->enqueue(*g_, *sub_);
    if (o_.profile) pGen_.end(sc_);
    if (d_bounds_) {
        // the scheduler's reduction and inference, one workgroup
        PolicyArgs pa{};
        pa.activeNodes = sub_->d_activeNodes;
        pa.anp = sub_->d_activeNodesPointer;
        pa.totals = gen_->deviceTotals();
        pa.hist = gen_->scanScratch().ctl->hist;
        pa.updatesNow = &fr_.ctl->updates;
        pa.state = d_pstate_;
        pa.bounds = d_bounds_;
        pa.P = (int)bounds_.size() - 1;
        pa.nodePointer = g_->d_nodePointer;
        pa.params = onPlain_ ? nullptr : d_params_;
        pa.V = g_->num_nodes;
        pa.E = g_->num_edges;
        pa.bufEdges = maxPart_;
        pa.iteration = st_.globalIterations + 1;
        pa.pageRank = isPR ? 1u : 0u;
        pa.out = reinterpret_cast<PolicyOut *>(d_pol_);
        pa.rstats = reinterpret_cast<ull *>(d_pol_ + sizeof(PolicyOut));
        launch_policy(pa, sc_);
        SUBWAY_CHECK(hipMemcpyAsync(h_pol_, d_pol_, polBytes_, hipMemcpyDeviceToHost, sc_));
    }
    // error words of every bounded wait so far; the partitioner synchronises
    // the stream before they are read
    errorCopies();
    part_->partition(*sub_, gen_->deviceTotals(), maxPart_);
    checkWaits();
    sub_->numActiveNodes = (uint)part_->numActiveNodes;
    sub_->numActiveEdges = part_->numActiveEdges;
}

// A bounded device wait that gave up leaves partial results: the run fails
// and the engine refuses further runs.
template <class E, class G>
void Engine<E, G>::checkWaits()
{
    static const char *where[3] = {"frontier scan", "generator", "rounds kernel"};
    for (int i = 0; i < 3; i++) {
        if (h_err_[i]) {
            broken_ = true;
            const std::string what = h_err_[i] == dev::WAIT_DATA
                                         ? std::string("device data check failed in the ") + where[i] + " (" +
                                               wait_error_name(h_err_[i]) + ")"
                                         : std::string("device wait gave up in the ") + where[i] + " (" +
                                               wait_error_name(h_err_[i]) + ", " +
                                               std::to_string(fr_.spinLimit) + " polls)";
            std::string diag;
            if (i == 2 && h_err_[i] == dev::WAIT_BARRIER) {
                dev::PersistCtl pc;
                SUBWAY_CHECK(hipMemcpy(&pc, pc_, sizeof(pc), hipMemcpyDeviceToHost));
                diag = " [round " + std::to_string(pc.diag[0] >> 1) + (pc.diag[0] & 1 ? " advance" : " frontier") +
                       " barrier: " + std::to_string(pc.diag[1]) + " members of " + std::to_string(pc.diag[4]) +
                       " workgroups, group counter " + std::to_string(pc.diag[2]) + ", top counter " +
                       std::to_string(pc.diag[3]) + "]";
            }
            throw Error(ERR_INTERNAL, what + diag + " at global iteration " + std::to_string(st_.globalIterations) +
                                          "; the session must be rebuilt");
        }
    }
}

template <class E, class G>
void Engine<E, G>::setupRanges()
{
    const ull n = g_->num_nodes;
    const ull *np = g_->nodePointer;
    // ranges of about an eighth of a buffer: the schedule decides per range,
    // consecutive whole ranges share a unit
    const ull rangeEdges = std::max<ull>(maxPart_ / 8, (ull)ADV_CHUNK);
    bounds_.assign(1, 0);
    ull v = 0;
    while (v < n) {
        // largest t with np[t] - np[v] <= rangeEdges
        ull lo = v, hi = n;
        while (lo < hi) {
            ull mid = lo + (hi - lo + 1) / 2;
            if (np[mid] - np[v] <= rangeEdges) lo = mid;
            else hi = mid - 1;
        }
        v = (lo == v) ? v + 1 : lo;  // a vertex larger than a buffer is a range of its own
        bounds_.push_back((uint)v);
    }
    const size_t P = bounds_.size() - 1;
    SUBWAY_CHECK(hipMalloc(&d_bounds_, bounds_.size() * sizeof(uint32_t)));
    SUBWAY_CHECK(hipMemcpy(d_bounds_, bounds_.data(), bounds_.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    polBytes_ = sizeof(PolicyOut) + P * 4 * sizeof(ull);
    SUBWAY_CHECK(hipMalloc(&d_pol_, polBytes_));
    SUBWAY_CHECK(hipHostMalloc((void **)&h_pol_, polBytes_, 0));
    h_rstats_ = reinterpret_cast<ull *>(h_pol_ + sizeof(PolicyOut));
    SUBWAY_CHECK(hipMalloc(&d_pstate_, sizeof(PolicyState)));
    SUBWAY_CHECK(hipMemset(d_pstate_, 0, sizeof(PolicyState)));
    SUBWAY_CHECK(hipDeviceSynchronize());
}

template <class E, class G>
void Engine<E, G>::pushHubUnits(uint part, uint v, ull fromNode, ull fromEdge, ull edges)
{
    const ull base = g_->nodePointer[v];
    for (ull c = 0; c * maxPart_ < edges; c++) {
        Unit h{};
        h.part = part;
        h.hub = true;
        h.chunk = c;
        h.dma = true;
        h.hostEdge = base + c * maxPart_;
        h.count = std::min(maxPart_, edges - c * maxPart_);
        setItems(h, fromNode, 1, fromEdge);
        units_.push_back(h);
    }
}

// A unit that moves every edge of the ids [idLo, idHi) by DMA moves them in
// the block layout when the graph holds it, the rows fit a buffer with two
// words of slack (the decoder reads a field's second word unconditionally)
// and the blocks are smaller than the plain edges.
template <class E, class G>
void Engine<E, G>::markBlocked(Unit &u) const
{
    u.blocked = false;
    if (!blkOk_ || !u.dma || u.hub || u.count == 0 || u.idHi <= u.idLo) return;
    const ull lo = g_->csr->blkPtr[u.idLo], words = g_->csr->blkPtr[u.idHi] - lo;
    if ((words + 2) * 4 > bufBytes_ || words + 2 >= (1ull << 32)) return;
    if (words * 4 >= u.count * (ull)eb_) return;
    u.blocked = true;
    u.blkLo = lo;
    u.blkWords = words;
}

// Units gathered from the active list (not a contiguous slice) move their
// rows in the block layout when that takes fewer bytes than their plain
// edges and fits a buffer: the row word counts of the active list are
// scanned once per global iteration, and the scan's values at each unit's
// item bounds come back to size the unit.
template <class E, class G>
void Engine<E, G>::markGatheredBlocked()
{
    if (!d_wpre_ || o_.transfer == TRANSFER_HOST) return;
    std::vector<size_t> cand;
    for (size_t i = 0; i < units_.size(); i++) {
        const Unit &u = units_[i];
        if (!u.dma && !u.hub && !u.blocked && u.count > 0 && u.nodes && u.kend > u.kbeg) cand.push_back(i);
    }
    if (cand.empty()) return;
    if (edgesMapped_) {
        // the plain edges are gathered by the GPU as well: the scan and its
        // read-back pay only when the units are large (16 MiB, or an eighth
        // of a buffer when buffers are smaller than 128 MiB)
        ull plain = 0;
        for (size_t i : cand) plain += units_[i].count * (ull)eb_;
        if (plain < std::min<ull>(kGatherBlockedMinBytes, bufBytes_ / 8)) return;
    }
    const ull A = sub_->numActiveNodes;
    if (!wpreValid_) {
        launch_row_words(A, sub_->d_activeNodes, d_blk_, d_wpre_, rw_, sc_);
        wpreValid_ = true;
    }
    if (2 * cand.size() > wbCap_) {
        if (h_wb_) SUBWAY_CHECK(hipHostFree(h_wb_));
        wbCap_ = std::max<size_t>(2 * cand.size(), 64);
        SUBWAY_CHECK(hipHostMalloc((void **)&h_wb_, wbCap_ * sizeof(ull), 0));
    }
    for (size_t c = 0; c < cand.size(); c++) {
        const Unit &u = units_[cand[c]];
        if (u.kend > A) throw Error(ERR_INTERNAL, "gathered unit past the active list");
        SUBWAY_CHECK(hipMemcpyAsync(&h_wb_[2 * c], d_wpre_ + u.kbeg, sizeof(ull), hipMemcpyDeviceToHost, sc_));
        SUBWAY_CHECK(hipMemcpyAsync(&h_wb_[2 * c + 1], d_wpre_ + u.kend, sizeof(ull), hipMemcpyDeviceToHost, sc_));
    }
    SUBWAY_CHECK(hipStreamSynchronize(sc_));
    for (size_t c = 0; c < cand.size(); c++) {
        Unit &u = units_[cand[c]];
        const ull lo = h_wb_[2 * c], hi = h_wb_[2 * c + 1];
        if (hi < lo) throw Error(ERR_INTERNAL, "row word scan is not monotone");
        const ull words = hi - lo;
        if ((words + 2) * 4 > bufBytes_ || words + 2 >= (1ull << 32) || words * 4 >= u.count * (ull)eb_) continue;
        u.blocked = true;
        u.blkLo = lo;
        u.blkWords = words;
    }
}

// Learned policy, per global iteration, after the policy kernel: the
// network's heads give the next iteration's active-edge fraction, the
// whole-range density threshold and the inner-round cap.  The plain policy is
// the fallback and the judge: a prediction more than 2x off the active edges
// the iteration then shows is a misprediction, and once more than half of at
// least eight judged predictions missed, the rest of the run goes to the
// plain schedule (iterations with under 1e-4 of the edges active, predicted
// and actual, are not judged).  Without parameters the options' fixed theta
// and cap apply.
// The schedules of the training sweep: whole-range density thresholds and
// inner-round caps.  Head 1 is theta; head 2 is log2(cap) / 16.
static const double kThetas[4] = {0.25, 0.5, 0.8, 2.0};
static const int kCaps[4] = {1, 8, 64, 4096};

static double snap_theta(float h)
{
    double best = kThetas[0];
    for (double t : kThetas)
        if (std::fabs((double)h - t) < std::fabs((double)h - best)) best = t;
    return best;
}

static int snap_cap(float h)
{
    const double l = 16.0 * std::min(std::max((double)h, 0.0), 1.0);
    int best = kCaps[0];
    for (int c : kCaps)
        if (std::fabs(l - std::log2((double)c)) < std::fabs(l - std::log2((double)best))) best = c;
    return best;
}

template <class E, class G>
void Engine<E, G>::applyPolicy(IterRecord &rec)
{
    const PolicyOut &po = *reinterpret_cast<const PolicyOut *>(h_pol_);
    std::memcpy(rec.features, po.feat, sizeof(rec.features));
    std::memcpy(rec.heads, po.head, sizeof(rec.heads));
    if (d_params_ && !onPlain_) {
        const float actual = po.feat[2];
        // iterations with under 1e-4 of the edges active (predicted and
        // actual) are not judged: their schedule hardly matters
        if (predEdges_ >= 0.f && (predEdges_ >= 1e-4f || actual >= 1e-4f)) {
            const float r = (predEdges_ + 1e-6f) / (actual + 1e-6f);
            judged_++;
            if (r > 2.f || r < 0.5f) st_.mispredictions++;
            // more than half of at least eight judged predictions missed
            if (judged_ >= 8 && 2 * st_.mispredictions > judged_) {
                onPlain_ = true;
                st_.fallbackIteration = st_.globalIterations;
            }
        }
        if (!onPlain_) {
            // head 0: log2 of the next iteration's active-edge fraction, / 32
            predEdges_ = (float)std::exp2(32.0 * std::min(std::max((double)po.head[0], -1.0), 0.0));
            // decisions snap to the schedules the network was trained on
            theta_ = snap_theta(po.head[1]);
            roundsCap_ = std::min(snap_cap(po.head[2]), o_.maxRounds);
        }
    }
    if (onPlain_) {
        theta_ = 2.0;   // never whole
        roundsCap_ = o_.maxRounds;
    }
    rec.theta = (float)theta_;
    rec.roundsCap = (uint)roundsCap_;
    rec.onPlain = onPlain_ ? 1u : 0u;
}

// Learned schedule over static id ranges: a range whose active-edge density
// (active edges / range edges) reaches theta_ moves whole by DMA and its
// rounds cover every vertex of the range; consecutive whole ranges share a
// unit up to one buffer.  The other ranges' active slices are packed into
// compacted units of at most one buffer (DMA when they form one fully active
// contiguous run, else gathered).  Vertices larger than a buffer stream in
// chunks.  Before all of that, ranges whose whole edge slice a buffer still
// holds from an earlier unit of the run are visited in place when their
// active edges come to kResidentMinBytes or more: those units go first, each
// at the position the iteration's rotation of the unit -> buffer map gives
// its buffer, and need no transfer.  Below that the transfer saved (about
// 150 us of the host link) does not pay for the visit's own persistent
// launch and its scans of every id of the ranges in each round.
static const ull kResidentMinBytes = 8ull << 20;
template <class E, class G>
void Engine<E, G>::buildUnitsLearned()
{
    units_.clear();
    bufShift_ = 0;
    const ull *np = g_->nodePointer;
    const size_t P = bounds_.size() - 1;
    const int nb = sub_->numBuffers;
    std::vector<char> done(P, 0);
    Unit res[2] = {};
    bool hasRes[2] = {false, false};
    if (!o_.noResident) {
        const ull minBytes = o_.residentMinBytes ? (ull)o_.residentMinBytes : kResidentMinBytes;
        const ResidentChoice rc = choose_resident_visits(
            bounds_.data(), P, [&](size_t i) { return h_rstats_[4 * i + 2]; }, bufIdLo_, bufIdHi_, nb, minBytes,
            (ull)eb_, done.data());
        st_.residentOverlaps += rc.overlaps;
        for (int b = 0; b < nb; b++) {
            if (!rc.visit[b]) continue;
            const uint lo = bufIdLo_[b], hi = bufIdHi_[b];
            Unit &u = res[b];
            u.dma = true;
            u.hostEdge = np[lo];
            u.count = np[hi] - np[lo];
            u.whole = true;
            u.n = hi - lo;
            u.nodes = nullptr;
            u.ptr = g_->d_nodePointer + lo;
            u.fromEdge = np[lo];
            u.idBase = lo;
            u.idLo = lo;
            u.idHi = hi;
            markBlocked(u);
            hasRes[b] = true;
            st_.residentVisits++;
        }
    }
    // resident units first; a lone one in buffer 1 rotates the map by one
    if (hasRes[0]) units_.push_back(res[0]);
    if (nb > 1 && hasRes[1]) {
        if (!hasRes[0]) bufShift_ = 1;
        units_.push_back(res[1]);
    }
    uint part = (uint)units_.size();
    for (uint k = 0; k < part; k++) units_[k].part = k;
    // open compacted group
    bool open = false;
    ull gLo = 0, gHi = 0, gEdges = 0, gBase = 0;
    bool gFull = true;
    uint gFirstRange = 0, gLastRange = 0;
    // open whole group: vertex ids [wB0, wB1), at most wCap edges, its last
    // range wLast
    bool wOpen = false;
    uint wB0 = 0, wB1 = 0;
    ull wEdges = 0, wCap = 0;
    size_t wLast = 0;
    // edges of the whole ranges from range i on (the ranges this iteration
    // moves whole), for the geometric tail of whole units, and their cost
    // (edges + kVertexCost * vertices) for the pieces of low-degree units
    std::vector<ull> wholeLeft(P + 1, 0), wholeCostLeft(P + 1, 0);
    for (size_t i = P; i-- > 0;) {
        const ull lo = h_rstats_[4 * i], hi = h_rstats_[4 * i + 1], ea = h_rstats_[4 * i + 2];
        const ull re = np[bounds_[i + 1]] - np[bounds_[i]];
        const bool hub = (bounds_[i + 1] - bounds_[i] == 1) && re > maxPart_;
        const bool whole = !done[i] && hi > lo && !hub && re > 0 && (double)ea >= theta_ * (double)re;
        wholeLeft[i] = wholeLeft[i + 1] + (whole ? re : 0);
        wholeCostLeft[i] = wholeCostLeft[i + 1] + (whole ? re + kVertexCost * (ull)(bounds_[i + 1] - bounds_[i]) : 0);
    }
    auto flush = [&]() {
        if (!open) return;
        Unit u{};
        u.part = part++;
        u.count = gEdges;
        u.kbeg = gLo;
        u.kend = gHi;
        u.edgeBase = gBase;
        u.dma = gFull || gEdges == 0;
        u.hostEdge = np[bounds_[gFirstRange]];
        if (o_.transfer == TRANSFER_ZEROCOPY && gEdges) u.dma = false;
        if (u.dma && gFull) {
            u.idLo = bounds_[gFirstRange];
            u.idHi = bounds_[gLastRange + 1];
        }
        setItems(u, gLo, gHi - gLo, gBase);
        markBlocked(u);
        units_.push_back(u);
        open = false;
    };
    auto emitWhole = [&](uint a, uint b) {
        Unit u{};
        u.part = part++;
        u.dma = true;
        u.hostEdge = np[a];
        u.count = np[b] - np[a];
        u.whole = true;
        u.n = b - a;
        u.nodes = nullptr;
        u.ptr = g_->d_nodePointer + a;
        u.fromEdge = np[a];
        u.idBase = a;
        u.idLo = a;
        u.idHi = b;
        markBlocked(u);
        units_.push_back(u);
    };
    std::vector<uint32_t> pieceEnds;
    auto flushWhole = [&]() {
        if (!wOpen) return;
        wOpen = false;
        if (kVertexCost * (ull)(wB1 - wB0) <= wEdges) {
            emitWhole(wB0, wB1);
            return;
        }
        // low-degree rows: pieces by cost, shrinking towards the iteration's end
        pieceEnds.clear();
        split_by_cost(wB0, wB1, np, wholeCostLeft[wLast + 1], kTailMinEdges, pieceEnds);
        uint a = wB0;
        for (uint32_t e : pieceEnds) {
            emitWhole(a, e);
            a = e;
        }
    };
    for (size_t i = 0; i < P; i++) {
        if (done[i]) {
            // visited in place: a compacted group must not span its active
            // vertices (they lie between the group's slices in the active list)
            flush();
            flushWhole();
            continue;
        }
        const ull lo = h_rstats_[4 * i], hi = h_rstats_[4 * i + 1], ea = h_rstats_[4 * i + 2];
        const ull base = h_rstats_[4 * i + 3];
        const uint b0 = bounds_[i], b1 = bounds_[i + 1];
        const ull re = np[b1] - np[b0];
        if (hi == lo) {
            // nothing active: the open groups stay open, but a compacted group
            // no longer covers one contiguous run of edges
            if (open && re > 0) gFull = false;
            flushWhole();
            continue;
        }
        const bool hub = (b1 - b0 == 1) && re > maxPart_;
        if (hub) {
            flush();
            flushWhole();
            pushHubUnits(part++, b0, lo, base, re);
            continue;
        }
        if (re > 0 && (double)ea >= theta_ * (double)re) {
            flush();
            if (wOpen && wB1 == b0 && wEdges + re <= wCap) {
                wB1 = b1;
                wEdges += re;
            } else {
                flushWhole();
                wOpen = true;
                wB0 = b0;
                wB1 = b1;
                wEdges = re;
                wCap = std::min<ull>(maxPart_, whole_unit_cap(maxPart_, re, wholeLeft[i]));
            }
            wLast = i;
            continue;
        }
        flushWhole();
        const bool full = (hi - lo) == (ull)(b1 - b0);
        if (open && gEdges + ea > maxPart_) flush();
        if (!open) {
            open = true;
            gLo = lo;
            gHi = hi;
            gEdges = ea;
            gBase = base;
            gFull = full;
            gFirstRange = (uint)i;
            gLastRange = (uint)i;
        } else {
            // the active slices of consecutive ranges are adjacent in the active list
            gFull = gFull && full;
            gHi = hi;
            gEdges += ea;
            gLastRange = (uint)i;
        }
    }
    flush();
    flushWhole();
}

template <class E, class G>
void Engine<E, G>::buildUnits()
{
    bufShift_ = 0;
    if (resident_) {
        // the whole graph as one unit: moved by the first iteration's DMA,
        // then read in place; its rounds cover every id
        units_.clear();
        if (part_->numPartitions == 0) return;
        Unit u{};
        u.part = 0;
        u.dma = true;
        u.hostEdge = 0;
        u.count = g_->num_edges;
        u.whole = true;
        u.n = g_->num_nodes;
        u.nodes = nullptr;
        u.ptr = g_->d_nodePointer;
        u.fromEdge = 0;
        u.idBase = 0;
        u.idLo = 0;
        u.idHi = g_->num_nodes;
        markBlocked(u);
        units_.push_back(u);
        return;
    }
    if (d_bounds_ && !onPlain_) {
        buildUnitsLearned();
        return;
    }
    units_.clear();
    const Partitioner<E> &p = *part_;
    for (uint i = 0; i < p.numPartitions; i++) {
        Unit u{};
        u.part = i;
        const ull cnt = p.partitionNodeSize[i], edges = p.partitionEdgeSize[i];
        if (edges == 0) {
            // nothing to move; labels of its vertices are still processed
            u.dma = true;
            u.count = 0;
            setItems(u, p.fromNode[i], cnt, p.fromEdge[i]);
            units_.push_back(u);
            continue;
        }
        if (cnt == 1 && edges > maxPart_) {
            const uint v = p.firstVertex[i];
            const ull base = g_->nodePointer[v];
            for (ull c = 0; c * maxPart_ < edges; c++) {
                Unit h{};
                h.part = i;
                h.hub = true;
                h.chunk = c;
                h.dma = true;
                h.hostEdge = base + c * maxPart_;
                h.count = std::min(maxPart_, edges - c * maxPart_);
                setItems(h, p.fromNode[i], cnt, p.fromEdge[i]);
                units_.push_back(h);
            }
            continue;
        }
        const bool contiguous = (ull)p.lastVertex[i] - (ull)p.firstVertex[i] + 1 == cnt;
        if (contiguous && o_.transfer != TRANSFER_ZEROCOPY) {
            u.dma = true;
            u.hostEdge = g_->nodePointer[p.firstVertex[i]];
            u.idLo = p.firstVertex[i];
            u.idHi = p.lastVertex[i] + 1;
        } else {
            u.dma = false;
        }
        u.count = edges;
        u.kbeg = p.fromNode[i];
        u.kend = p.fromNode[i] + cnt;
        u.edgeBase = p.fromEdge[i];
        setItems(u, p.fromNode[i], cnt, p.fromEdge[i]);
        markBlocked(u);
        units_.push_back(u);
    }
}

template <class E, class G>
void Engine<E, G>::issue(size_t idx)
{
    const Unit &u = units_[idx];
    const int b = (int)((idx + bufShift_) % (size_t)sub_->numBuffers);
    SUBWAY_CHECK(hipStreamWaitEvent(sx_, evFree_[b], 0));
    if (resident_ && residentLoaded_) {
        SUBWAY_CHECK(hipEventRecord(evLoaded_[b], sx_));   // already in buffer 0
        return;
    }
    if (resident_) residentLoaded_ = true;
    if (u.dma && u.count && bufCount_[b] == u.count && bufHostEdge_[b] == u.hostEdge &&
        bufBlocked_[b] == u.blocked) {
        SUBWAY_CHECK(hipEventRecord(evLoaded_[b], sx_));   // the slice is still in the buffer
        return;
    }
    bufCount_[b] = u.dma ? u.count : 0;
    bufHostEdge_[b] = u.hostEdge;
    bufBlocked_[b] = u.blocked;
    const bool slice = u.dma && u.count && u.idHi > u.idLo;
    bufIdLo_[b] = slice ? u.idLo : 0u;
    bufIdHi_[b] = slice ? u.idHi : 0u;
    if (u.count) {
        if (o_.profile) pXfer_.begin(sx_);
        char *dst = (char *)sub_->d_buffers[b];
        const size_t bytes = (size_t)u.count * (size_t)eb_;
        const bool split = fmt_ == EDGE_SPLIT;
        // split format: destinations at the buffer's start, weights at byte 4 * capacity
        char *dstW = split ? dst + 4ull * maxPart_ : nullptr;
        if (u.blocked && !u.dma) {
            // the rows of the unit's active vertices, read by the GPU from
            // the mapped block words, packed in active-list order
            const size_t bb = (size_t)u.blkWords * 4;
            launch_gather(4, sub_->d_activeNodes, d_wpre_, u.kbeg, u.kend, d_blk_, blkDev_, dst, u.blkLo, u.blkWords,
                          sx_);
            st_.h2dGatherBytes += bb;
            st_.h2dBlockedBytes += bb;
            st_.blockedEdges += u.count;
            st_.blockedUnits++;
            st_.h2dBytes += bb;
        } else if (u.blocked) {
            const size_t bb = (size_t)u.blkWords * 4;
            const hipError_t ce = hipMemcpyAsync(dst, g_->csr->blk + u.blkLo, bb, hipMemcpyHostToDevice, sx_);
            if (ce != hipSuccess) {
                (void)hipGetLastError();
                if (!blkDiag_) {
                    blkDiag_ = true;
                    size_t sz = 0;
                    const hipError_t ie = hipMemPtrGetInfo(g_->csr->blk, &sz);
                    hipPointerAttribute_t at{};
                    const hipError_t pe = hipPointerGetAttributes(&at, g_->csr->blk + u.blkLo);
                    std::fprintf(stderr,
                                 "subway: block copy refused (%s): words [%llu, +%llu) of %llu; allocation %s size %zu; "
                                 "pointer %s type %d host %p device %p\n",
                                 hipGetErrorString(ce), (unsigned long long)u.blkLo, (unsigned long long)u.blkWords,
                                 (unsigned long long)g_->csr->blkWords, hipGetErrorString(ie), sz,
                                 hipGetErrorString(pe), (int)at.type, at.hostPointer, at.devicePointer);
                    (void)hipGetLastError();
                }
                SUBWAY_CHECK(ce);
            }
            st_.h2dDmaBytes += bb;
            st_.h2dBlockedBytes += bb;
            st_.blockedEdges += u.count;
            st_.blockedUnits++;
            st_.h2dBytes += bb;
        } else if (u.dma) {
            const HostCsr<E> &c = *g_->csr;
            if (split && !edgesMapped_ && c.tailDst && u.hostEdge >= c.tailEdge) {
                // the pinned copy of the low-degree id suffix
                SUBWAY_CHECK(hipMemcpyAsync(dst, c.tailDst + (u.hostEdge - c.tailEdge), 4ull * u.count,
                                            hipMemcpyHostToDevice, sx_));
                SUBWAY_CHECK(hipMemcpyAsync(dstW, c.tailW + (u.hostEdge - c.tailEdge), u.count,
                                            hipMemcpyHostToDevice, sx_));
            } else if (split && !edgesMapped_) {
                // split arrays in ordinary memory: through the pinned staging slots
                stagedCopy((const char *)(c.splitDst + u.hostEdge), dst, 4ull * u.count);
                stagedCopy((const char *)(c.splitW + u.hostEdge), dstW, u.count);
            } else if (split) {
                SUBWAY_CHECK(hipMemcpyAsync(dst, g_->csr->splitDst + u.hostEdge, 4ull * u.count,
                                            hipMemcpyHostToDevice, sx_));
                SUBWAY_CHECK(hipMemcpyAsync(dstW, g_->csr->splitW + u.hostEdge, u.count, hipMemcpyHostToDevice,
                                            sx_));
            } else {
                SUBWAY_CHECK(hipMemcpyAsync(dst, hostEdges() + u.hostEdge * (ull)eb_, bytes,
                                            hipMemcpyHostToDevice, sx_));
            }
            st_.h2dDmaBytes += bytes;
        } else if (o_.transfer == TRANSFER_HOST || !edgesMapped_) {
            if (split) {
                hostGather(u, (const char *)g_->csr->splitDst, 4, dst);
                hostGather(u, (const char *)g_->csr->splitW, 1, dstW);
            } else {
                hostGather(u, hostEdges(), (size_t)eb_, dst);
            }
            st_.h2dDmaBytes += bytes;
        } else {
            if (split) {
                launch_gather(4, sub_->d_activeNodes, sub_->d_activeNodesPointer, u.kbeg, u.kend, g_->d_nodePointer,
                              g_->csr->d_splitDst, dst, u.edgeBase, u.count, sx_);
                launch_gather(1, sub_->d_activeNodes, sub_->d_activeNodesPointer, u.kbeg, u.kend, g_->d_nodePointer,
                              g_->csr->d_splitW, dstW, u.edgeBase, u.count, sx_);
            } else {
                const void *mapped =
                    fmt_ == EDGE_PACKED ? (const void *)g_->csr->d_packed : (const void *)g_->d_edgeList;
                launch_gather(eb_, sub_->d_activeNodes, sub_->d_activeNodesPointer, u.kbeg, u.kend,
                              g_->d_nodePointer, mapped, dst, u.edgeBase, u.count, sx_);
            }
            st_.h2dGatherBytes += bytes;
        }
        if (!u.blocked) st_.h2dBytes += bytes;
        if (o_.profile) pXfer_.end(sx_);
    }
    SUBWAY_CHECK(hipEventRecord(evLoaded_[b], sx_));
}

// Host-thread gather (the reference's dynamic<E>, subgraph_generator.cu:63-97,
// without its tail bug): the partition's active ids are mirrored once per
// global iteration; host threads compute the unit's row offsets (chunk sums,
// then a prefix inside each chunk), then copy the rows chunk by chunk into
// one of two pinned staging slots of kStageBytes, each chunk followed by its
// DMA to dstDev on the copy stream: the threads fill one slot while the other
// one's copy runs, and the pinned memory a session holds stays small.
template <class E, class G>
void Engine<E, G>::hostGather(const Unit &u, const char *src, size_t esz, char *dstDev)
{
    if (!hostIdsValid_) {
        const ull A = sub_->numActiveNodes;
        SUBWAY_CHECK(hipMemcpyAsync(pinnedIds(A), sub_->d_activeNodes, A * sizeof(uint), hipMemcpyDeviceToHost, sc_));
        SUBWAY_CHECK(hipStreamSynchronize(sc_));
        hostIdsValid_ = true;
    }
    const uint *ids = h_ids_ + u.kbeg;
    const ull cnt = u.kend - u.kbeg;
    const ull *np = g_->nodePointer;
    const int T = std::max(1, std::min(host_threads(), 64));
    std::vector<ull> &pre = gatherPre_;
    pre.resize(cnt + 1);
    const ull nch = std::max<ull>(1, std::min<ull>(cnt, (ull)T * 8));
    std::vector<ull> chunkEdges(nch + 1, 0);
    parallel_for(nch, [&](ull c0, ull c1, int) {
        for (ull c = c0; c < c1; c++) {
            ull v0 = cnt * c / nch, v1 = cnt * (c + 1) / nch, e = 0;
            for (ull i = v0; i < v1; i++) e += np[ids[i] + 1] - np[ids[i]];
            chunkEdges[c + 1] = e;
        }
    }, T);
    for (ull c = 0; c < nch; c++) chunkEdges[c + 1] += chunkEdges[c];
    if (chunkEdges[nch] != u.count) throw Error(ERR_INTERNAL, "host gather: edge count mismatch");
    parallel_for(nch, [&](ull c0, ull c1, int) {
        for (ull c = c0; c < c1; c++) {
            ull v0 = cnt * c / nch, v1 = cnt * (c + 1) / nch, at = chunkEdges[c];
            for (ull i = v0; i < v1; i++) {
                pre[i] = at;
                at += np[ids[i] + 1] - np[ids[i]];
            }
        }
    }, T);
    pre[cnt] = u.count;
    const ull per = kStageBytes / esz;
    for (ull a = 0; a < u.count; a += per) {
        const ull e1 = std::min<ull>(u.count, a + per);
        int s = 0;
        char *stage = stageSlot(s);
        parallel_for(e1 - a, [&](ull x0, ull x1, int) {
            ull e = a + x0;
            const ull end = a + x1;
            // the row holding edge e: the last one that starts at or before it
            ull k = (ull)(std::upper_bound(pre.begin(), pre.begin() + cnt + 1, e) - pre.begin()) - 1;
            while (e < end) {
                const ull take = std::min(pre[k + 1] - e, end - e);
                if (take) std::memcpy(stage + (e - a) * esz, src + (np[ids[k]] + (e - pre[k])) * esz, take * esz);
                e += take;
                k++;
            }
        }, T);
        SUBWAY_CHECK(hipMemcpyAsync(dstDev + a * esz, stage, (e1 - a) * esz, hipMemcpyHostToDevice, sx_));
        SUBWAY_CHECK(hipEventRecord(evStaged_[s], sx_));
    }
}

// The next of the two pinned staging slots, once its previous copy finished.
template <class E, class G>
char *Engine<E, G>::stageSlot(int &s)
{
    if (!stage_[0]) {
        for (int k = 0; k < 2; k++) SUBWAY_CHECK(hipHostMalloc((void **)&stage_[k], kStageBytes, 0));
    }
    s = stageSlot_;
    stageSlot_ ^= 1;
    SUBWAY_CHECK(hipEventSynchronize(evStaged_[s]));
    return stage_[s];
}

template <class E, class G>
void Engine<E, G>::stagedCopy(const char *host, char *dstDev, size_t bytes)
{
    for (size_t a = 0; a < bytes; a += kStageBytes) {
        const size_t n = std::min<size_t>(kStageBytes, bytes - a);
        int s = 0;
        char *stage = stageSlot(s);
        parallel_for(n, [&](ull b, ull e, int) { std::memcpy(stage + b, host + a + b, e - b); });
        SUBWAY_CHECK(hipMemcpyAsync(dstDev + a, stage, n, hipMemcpyHostToDevice, sx_));
        SUBWAY_CHECK(hipEventRecord(evStaged_[s], sx_));
    }
}

template <class E, class G>
void Engine<E, G>::frontier(int first, uint8_t *labA, uint8_t *labB, const Unit &un)
{
    FrontierArgs a{};
    a.algo = o_.algo;
    a.n = un.n;
    a.nodes = un.nodes;
    a.ptr = un.ptr;
    a.fromEdge = un.fromEdge;
    a.idBase = un.idBase;
    a.labA = labA;
    a.labB = labB;
    if constexpr (isPR) {
        a.delta = d_resid_;
        a.sum = d_sum_;
        a.acc = o_.acc;
    } else {
        a.value = g_->d_value;
    }
    a.entries = entries_;
    a.chunkMap = chunkMap_;
    a.first = first;
    if (un.blocked) {
        // DMA units: row offsets by vertex id; gathered units: by item
        a.blkByItem = un.dma ? 0u : 1u;
        a.blk = un.dma ? d_blk_ : d_wpre_ + un.kbeg;
        a.blkBase = un.blkLo;
    }
    if (o_.profile) pFr_.begin(sc_);
    launch_frontier(a, fr_, sc_);
    if (o_.profile) pFr_.end(sc_);
}

template <class E, class G>
void Engine<E, G>::advance(uint8_t *labOut, ull w0, ull w1, ull shift, int buf, const Unit &un)
{
    AdvanceArgs a{};
    a.algo = o_.algo;
    a.edgeFmt = fmtOf(un);
    a.packCode = codeOf(un);
    a.entries = entries_;
    a.chunkMap = chunkMap_;
    a.edges = sub_->d_buffers[buf];
    if constexpr (isPR) {
        a.delta = d_resid_;
    } else {
        a.value = g_->d_value;
        a.labelOut = labOut;
    }
    a.w0 = w0;
    a.w1 = w1;
    a.shift = shift;
    a.grid = advGrid_;
    a.nv = g_->num_nodes;
    a.wlim = (uint32_t)std::min<ull>(bufBytes_ / 4, 0xFFFFFFFFull);
    if (o_.profile) pAdv_.begin(sc_);
    launch_advance(a, fr_.ctl, sc_);
    if (o_.profile) pAdv_.end(sc_);
}

template <class E, class G>
void Engine<E, G>::asyncRounds(const Unit &un, int buf)
{
    uint8_t *l1 = nullptr, *l2 = nullptr;
    if constexpr (!isPR) {
        l1 = (uint8_t *)g_->d_label1;
        l2 = (uint8_t *)g_->d_label2;
    }
    if (roundsCap_ == 1) {
        // a single round: one frontier + advance pair (the persistent launch
        // would add its membership and barrier costs for nothing); the labels
        // it sets are picked up by the next generate
        frontier(1, l1, l2, un);
        advance(l2, 0, ~0ull, 0, buf, un);
        return;
    }
    if (roundsGrid_ > 0) {
        // every round of the partition in one persistent launch; a partition
        // that has not converged after maxRounds keeps its labels and is
        // picked up by the next generate, so no host check is needed
        RoundsLaunch l{};
        l.n = un.n;
        l.nodes = un.nodes;
        l.ptr = un.ptr;
        l.fromEdge = un.fromEdge;
        l.idBase = un.idBase;
        l.l1 = l1;
        l.l2 = l2;
        if constexpr (isPR) {
            l.delta = d_resid_;
            l.sum = d_sum_;
            l.acc = o_.acc;
        } else {
            l.value = g_->d_value;
        }
        l.entries = entries_;
        l.chunkMap = chunkMap_;
        l.edges = sub_->d_buffers[buf];
        l.scan = &fr_;
        l.pc = pc_;
        l.maxRounds = (uint32_t)std::max(1, roundsCap_);
        l.roundBase = 0;
        EventPairs &ev = un.blocked ? pRoundsBlk_ : pRounds_;
        if (o_.profile) ev.begin(sc_);
        l.packCode = codeOf(un);
        if (un.blocked) {
            l.blkByItem = un.dma ? 0u : 1u;
            l.blk = un.dma ? d_blk_ : d_wpre_ + un.kbeg;
            l.blkBase = un.blkLo;
        }
        l.joinTicks = o_.joinWindow;
        l.nv = g_->num_nodes;
        l.wlim = (uint32_t)std::min<ull>(bufBytes_ / 4, 0xFFFFFFFFull);
        launch_rounds(o_.algo, fmtOf(un), l, roundsGrid_, sc_);
        if (o_.profile) ev.end(sc_);
        return;
    }
    ull r = 0;
    int batch = 2;
    for (;;) {
        for (int q = 0; q < batch; q++) {
            r++;
            const bool odd = (r & 1) != 0;
            if (r == 1) frontier(1, l1, l2, un);
            else frontier(0, odd ? l1 : l2, nullptr, un);
            advance(odd ? l2 : l1, 0, ~0ull, 0, buf, un);
        }
        SUBWAY_CHECK(hipMemcpyAsync(&h_ctl_->stop, &fr_.ctl->stop, sizeof(uint32_t), hipMemcpyDeviceToHost, sc_));
        SUBWAY_CHECK(hipMemcpyAsync(&h_err_[0], &fr_.ctl->error, sizeof(uint32_t), hipMemcpyDeviceToHost, sc_));
        SUBWAY_CHECK(hipStreamSynchronize(sc_));
        if (h_err_[0]) checkWaits();
        if (h_ctl_->stop) break;
        batch = std::min(batch * 2, std::max(1, o_.maxRoundBatch));
    }
}

// In-edges of the resident graph for the bottom-up steps of BFS: in-degree
// counts, their prefix and a scatter of every edge's source, by host threads
// with relaxed atomic counters (the order inside an in-list does not matter
// to BFS), moved to the device once per session.
template <class E, class G>
void Engine<E, G>::buildInEdges()
{
    const ull n = g_->num_nodes, m = g_->num_edges;
    const HostCsr<E> &c = *g_->csr;
    std::vector<ull> inPtr(n + 1, 0);
    std::vector<uint32_t> cnt(n, 0), adj(m);
    parallel_for(n, [&](uint64_t b, uint64_t e, int) {
        for (ull v = b; v < e; v++)
            for (ull k = c.ptr[v]; k < c.ptr[v + 1]; k++) __atomic_fetch_add(&cnt[c.edgeAt(k).end], 1u, __ATOMIC_RELAXED);
    });
    for (ull v = 0; v < n; v++) inPtr[v + 1] = inPtr[v] + cnt[v];
    parallel_for(n, [&](uint64_t b, uint64_t e, int) {
        for (ull v = b; v < e; v++) cnt[v] = 0;
    });
    parallel_for(n, [&](uint64_t b, uint64_t e, int) {
        for (ull v = b; v < e; v++)
            for (ull k = c.ptr[v]; k < c.ptr[v + 1]; k++) {
                const uint32_t d = c.edgeAt(k).end;
                adj[inPtr[d] + __atomic_fetch_add(&cnt[d], 1u, __ATOMIC_RELAXED)] = (uint32_t)v;
            }
    });
    SUBWAY_CHECK(hipMalloc(&d_inPtr_, (n + 1) * sizeof(ull)));
    SUBWAY_CHECK(hipMalloc(&d_inAdj_, (m ? m : 1) * sizeof(uint32_t)));
    SUBWAY_CHECK(hipMemcpy(d_inPtr_, inPtr.data(), (n + 1) * sizeof(ull), hipMemcpyHostToDevice));
    if (m) SUBWAY_CHECK(hipMemcpy(d_inAdj_, adj.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice));
}

// Direction-optimizing BFS on the resident graph (Beamer, Asanovic and
// Patterson, SC 2012), level by level: each step's frontier pass selects the
// labelled vertices (all at level `cur`), clears their labels and counts
// them and their out-edges; a top-down step then pushes along those edges
// (the advance), a bottom-up step has every unvisited vertex look for an
// in-neighbour at level `cur` instead.  Bottom-up when the frontier's edges
// exceed 1/14 of the edges of unvisited vertices, back to top-down when the
// frontier falls below 1/24 of the vertices (the published thresholds).
// Both steps label the next frontier; levels equal the push-only fixpoint.
template <class E, class G>
void Engine<E, G>::bfsDirectionOptimizing(const Unit &un, int buf)
{
    if constexpr (!isPR) {
        constexpr ull kAlpha = 14, kBeta = 24;
        const ull n = g_->num_nodes, m = g_->num_edges;
        uint8_t *la = (uint8_t *)g_->d_label1, *lb = (uint8_t *)g_->d_label2;
        ull explored = 0;
        bool pull = false;
        for (uint32_t cur = 0;; cur++) {
            // level 0: the source's label may sit in either array
            frontier(1, la, cur == 0 ? lb : nullptr, un);
            SUBWAY_CHECK(hipMemcpyAsync(h_ctl_, fr_.ctl, sizeof(ScanCtl), hipMemcpyDeviceToHost, sc_));
            SUBWAY_CHECK(hipStreamSynchronize(sc_));
            if (h_ctl_->error) {
                h_err_[0] = h_ctl_->error;
                checkWaits();
            }
            const ull nf = h_ctl_->count, mf = h_ctl_->total;
            if (nf == 0) break;
            const ull unexplored = m > explored ? m - explored : 0;
            if (!pull && mf * kAlpha > unexplored) pull = true;
            else if (pull && nf * kBeta < n) pull = false;
            explored += mf;
            if (pull) {
                launch_fill_u32(reinterpret_cast<uint32_t *>(&fr_.ctl->pull_found), 0u, 4, sc_);
                if (o_.profile) pAdv_.begin(sc_);
                launch_bfs_pull(n, d_inPtr_, d_inAdj_, g_->d_outDegree, g_->d_value, cur, lb, fr_.ctl, sc_);
                if (o_.profile) pAdv_.end(sc_);
                st_.pullRounds++;
                pullCounted_ = true;
            } else {
                advance(lb, 0, ~0ull, 0, buf, un);
            }
            std::swap(la, lb);
            if (pullCounted_) {
                SUBWAY_CHECK(hipMemcpyAsync(&h_ctl_->pull_found, &fr_.ctl->pull_found, sizeof(ull),
                                            hipMemcpyDeviceToHost, sc_));
                SUBWAY_CHECK(hipStreamSynchronize(sc_));
                st_.pullFound += h_ctl_->pull_found;
                pullCounted_ = false;
            }
        }
    } else {
        (void)un;
        (void)buf;
    }
}

template <class E, class G>
void Engine<E, G>::compute(size_t idx)
{
    const Unit &u = units_[idx];
    const int b = (int)((idx + bufShift_) % (size_t)sub_->numBuffers);
    SUBWAY_CHECK(hipStreamWaitEvent(sc_, evLoaded_[b], 0));
    uint8_t *l1 = nullptr, *l2 = nullptr;
    if constexpr (!isPR) {
        l1 = (uint8_t *)g_->d_label1;
        l2 = (uint8_t *)g_->d_label2;
    }
    const bool async = o_.mode == ASYNC;
    if (batchK_) {
        computeBatch(u, b);
    } else if (resident_) {
        // rounds to the fixpoint (or the round cap) on the resident graph, in
        // both modes: the label fixpoints are unique, PR stays within its bound
        if (d_inPtr_) bfsDirectionOptimizing(u, b);
        else asyncRounds(u, b);
    } else if (u.hub) {
        // one round over the oversized vertex, streamed chunk by chunk: the
        // frontier kernel runs once (clearing its label, fixing its value),
        // then each chunk is relaxed through a window of the work space
        if (u.chunk == 0) {
            if (isPR) frontier(1, nullptr, nullptr, u);
            else if (async) frontier(1, l1, l2, u);
            else frontier(1, l2, nullptr, u);
        }
        const ull w0 = u.chunk * maxPart_;
        advance(l2, w0, w0 + u.count, w0, b, u);
    } else if (!async) {
        if (isPR) frontier(1, nullptr, nullptr, u);
        else frontier(1, l2, nullptr, u);
        advance(l2, 0, ~0ull, 0, b, u);
    } else {
        asyncRounds(u, b);
    }
    SUBWAY_CHECK(hipEventRecord(evFree_[b], sc_));
    if (o_.verbose && (!u.hub || u.chunk == 0)) {
        std::cout << "Global Iteration " << st_.globalIterations << ", Partition " << u.part << ": " << u.n
                  << (u.whole ? " vertices (whole range), " : " vertices, ") << u.count << " edges\n";
    }
}

// Batched partition: all rounds for every source in one persistent launch;
// an oversized vertex gets one masked frontier and windowed advances.
template <class E, class G>
void Engine<E, G>::computeBatch(const Unit &u, int b)
{
    if (u.hub) {
        if (u.chunk == 0)
            launch_frontier_masks(u.n, u.nodes, u.ptr, u.fromEdge, d_m1_, d_m2_, entries_, d_entryMask_, chunkMap_,
                                  fr_, sc_, u.idBase);
        const ull w0 = u.chunk * maxPart_;
        launch_advance_batch_window(o_.algo, fmt_, packCode_, entries_, d_entryMask_, chunkMap_, sub_->d_buffers[b],
                                    d_dist_, d_m2_, w0, w0 + u.count, w0, fr_.ctl, advGrid_, sc_);
        return;
    }
    BatchLaunch l{};
    l.n = u.n;
    l.nodes = u.nodes;
    l.ptr = u.ptr;
    l.fromEdge = u.fromEdge;
    l.idBase = u.idBase;
    l.m1 = d_m1_;
    l.m2 = d_m2_;
    l.dist = d_dist_;
    l.entries = entries_;
    l.entryMask = d_entryMask_;
    l.chunkMap = chunkMap_;
    l.edges = sub_->d_buffers[b];
    l.scan = &fr_;
    l.pc = pc_;
    l.maxRounds = (uint32_t)std::max(1, roundsCap_);
    if (o_.profile) pRounds_.begin(sc_);
    l.packCode = packCode_;
    l.joinTicks = o_.joinWindow;
    launch_rounds_batch(o_.algo, fmt_, batchUnroll_, l, batchGrid_, sc_);
    if (o_.profile) pRounds_.end(sc_);
}

template <class E, class G>
ull Engine<E, G>::deviceRounds()
{
    SUBWAY_CHECK(hipMemcpyAsync(h_ctl_, fr_.ctl, sizeof(ScanCtl), hipMemcpyDeviceToHost, sc_));
    SUBWAY_CHECK(hipStreamSynchronize(sc_));
    return h_ctl_->rounds;
}

template <class E, class G>
void Engine<E, G>::resetRun()
{
    if (broken_) throw Error(ERR_INTERNAL, "a device wait of an earlier run gave up; the session must be rebuilt");
    if (early_) {
        // an earlier run ended by an error after its early copy: it must
        // land before this run writes the result array
        SUBWAY_CHECK(hipEventSynchronize(evBack_));
        early_ = false;
    }
    st_ = RunStats();
    residentLoaded_ = false;
    bufCount_[0] = bufCount_[1] = 0;   // every run moves its edges at least once
    bufIdLo_[0] = bufIdLo_[1] = bufIdHi_[0] = bufIdHi_[1] = 0;
    bufShift_ = 0;
    roundsCap_ = o_.maxRounds;
    theta_ = o_.theta;
    onPlain_ = false;
    predEdges_ = -1.f;
    judged_ = 0;
    if (d_pstate_) {
        // updates counted from here on belong to this run
        SUBWAY_CHECK(hipMemcpyAsync(&d_pstate_->prevUpdates, &fr_.ctl->updates, sizeof(ull),
                                    hipMemcpyDeviceToDevice, sc_));
    }
    if (pc_) launch_fill_u32(&pc_->min_members, 0xFFFFFFFFu, 1, sc_);
    pAdv_.reset();
    pFr_.reset();
    pXfer_.reset();
    pGen_.reset();
    pRounds_.reset();
    pRoundsBlk_.reset();
}

// The global-iteration loop shared by single-source and batched runs; the
// first generate has been enqueued by the caller.  Returns with the device
// work of the last iteration drained from the copy stream.
template <class E, class G>
RunStats Engine<E, G>::loop(const std::chrono::steady_clock::time_point &t0)
{
    (void)t0;
    ull peakActiveEdges = 0;
    while (sub_->numActiveNodes > 0) {
        peakActiveEdges = std::max<ull>(peakActiveEdges, sub_->numActiveEdges);
        const auto ti = clk::now();
        st_.globalIterations++;
        IterRecord rec{};
        rec.iteration = st_.globalIterations;
        rec.activeNodes = sub_->numActiveNodes;
        rec.activeEdges = sub_->numActiveEdges;
        rec.partitions = part_->numPartitions;
        const ull h2d0 = st_.h2dBytes;
        if (d_bounds_) applyPolicy(rec);
        buildUnits();
        markGatheredBlocked();
        const size_t nu = units_.size();
        const size_t ahead = (size_t)sub_->numBuffers - 1;
        size_t issued = 0;
        // unit u is issued before its compute is enqueued, the units ahead
        // after it: an issue that works on the host (a gather or a staged
        // copy through the pinned slots) then runs while unit u computes
        for (size_t u = 0; u < nu; u++) {
            while (issued < nu && issued <= u) issue(issued++);
            compute(u);
            while (issued < nu && issued <= u + ahead) issue(issued++);
        }
        st_.partitions += (d_bounds_ && !onPlain_) ? (units_.empty() ? 0 : units_.back().part + 1)
                                                   : part_->numPartitions;
        rec.h2dBytes = st_.h2dBytes - h2d0;
        generateAndPartition();
        if (earlyRun_ && !early_ && sub_->numActiveNodes > 0 &&
            (o_.earlyReadback > 0 || sub_->numActiveEdges * 64 <= peakActiveEdges))
            startEarlyReadback();
        rec.msIteration = ms_since(ti);
        st_.trace.push_back(rec);
    }
    return st_;
}

// Device values (width u32 per vertex) into the host array, in the original
// vertex order: a plain copy, or on a renumbered graph a gather through perm
// into d_unperm_ first, in chunks of unpermRows_ vertices.
template <class E, class G>
void Engine<E, G>::copyValuesOut(const void *dValues, void *hValues, uint32_t width)
{
    const ull n = g_->num_nodes;
    if (!d_perm_) {
        SUBWAY_CHECK(hipMemcpyAsync(hValues, dValues, (size_t)n * width * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                    sc_));
        return;
    }
    for (ull v0 = 0; v0 < n; v0 += unpermRows_) {
        const ull cnt = std::min<ull>(unpermRows_, n - v0);
        if (v0) SUBWAY_CHECK(hipStreamSynchronize(sc_));   // d_unperm_ is free again
        launch_unpermute((const uint32_t *)dValues, d_perm_, v0, cnt, width, d_unperm_, sc_);
        SUBWAY_CHECK(hipMemcpyAsync((uint32_t *)hValues + v0 * width, d_unperm_, (size_t)cnt * width * sizeof(uint32_t),
                                    hipMemcpyDeviceToHost, sc_));
    }
}

template <class E, class G>
RunStats Engine<E, G>::run(uint source)
{
    if (batchK_ || o_.batch) throw Error(ERR_ARG, "this engine runs batched queries (runBatch)");
    if (o_.algo != PR && o_.algo != CC && source >= g_->num_nodes)
        throw Error(ERR_ARG, "source vertex " + std::to_string(source) + " out of range");
    // sources are original ids; a renumbered graph runs on internal ones
    src_ = (g_->csr->perm && source < g_->num_nodes) ? g_->csr->perm[source] : source;
    resetRun();
    initState();
    const uint n = g_->num_nodes;
    startStats();
    for (int b = 0; b < 2; b++) {
        SUBWAY_CHECK(hipEventRecord(evFree_[b], sc_));
        SUBWAY_CHECK(hipEventRecord(evStaged_[b], sx_));
    }
    SUBWAY_CHECK(hipDeviceSynchronize());

    const auto t0 = clk::now();
    generateAndPartition();
    if constexpr (!isPR) {
        if (o_.mode == ASYNC && o_.algo != CC) {
            // every vertex was active for the first generate; from now on only
            // the source is labelled (sssp-async.cu:59-64)
            launch_fill_u8((uint8_t *)g_->d_label1, 0, n, sc_);
            static const uint8_t one = 1;
            SUBWAY_CHECK(hipMemcpyAsync((uint8_t *)g_->d_label1 + src_, &one, 1, hipMemcpyHostToDevice, sc_));
        }
    }
    earlyRun_ = earlyOk_;
    loop(t0);
    earlyRun_ = false;
    if constexpr (isPR) launch_finish_pr(n, g_->d_value, d_sum_, g_->d_delta, d_resid_, sc_);
    finishValues();
    SUBWAY_CHECK(hipStreamSynchronize(sc_));
    st_.msTotal = ms_since(t0);
    finishStats();
    return st_;
}

// k sources (k <= 64) advance together: lane i of every distance row and bit
// i of every mask belong to sources[i].  Like the async apps, the first
// generate sees every vertex active (the whole graph moves once), then only
// the sources' bits are set.
template <class E, class G>
RunStats Engine<E, G>::runBatch(const uint *sources, int k)
{
    if (!o_.batch) throw Error(ERR_ARG, "the engine was not built for batched queries");
    if (k < 1 || k > BATCH_LANES) throw Error(ERR_ARG, "a batch holds 1..64 sources");
    const uint n = g_->num_nodes;
    for (int i = 0; i < k; i++)
        if (sources[i] >= n) throw Error(ERR_ARG, "batch source " + std::to_string(sources[i]) + " out of range");
    batchK_ = k;
    resetRun();
    for (int i = 0; i < k; i++) h_sources_[i] = g_->csr->perm ? g_->csr->perm[sources[i]] : sources[i];
    SUBWAY_CHECK(hipMemcpyAsync(d_sources_, h_sources_, (size_t)k * sizeof(uint32_t), hipMemcpyHostToDevice, sc_));
    launch_init_batch_state(o_.algo, n, d_dist_, d_m1_, d_m2_, 1, sc_);
    startStats();
    for (int b = 0; b < 2; b++) {
        SUBWAY_CHECK(hipEventRecord(evFree_[b], sc_));
        SUBWAY_CHECK(hipEventRecord(evStaged_[b], sx_));
    }
    if (!h_dist_) SUBWAY_CHECK(hipHostMalloc((void **)&h_dist_, (size_t)n * BATCH_LANES * sizeof(uint32_t), 0));
    SUBWAY_CHECK(hipDeviceSynchronize());

    const auto t0 = clk::now();
    generateAndPartition();
    launch_fill_u32((uint32_t *)d_m1_, 0, 2ull * n, sc_);
    launch_set_batch_sources(o_.algo, d_sources_, k, d_dist_, d_m1_, sc_);
    loop(t0);
    copyValuesOut(d_dist_, h_dist_, BATCH_LANES);
    SUBWAY_CHECK(hipStreamSynchronize(sc_));
    st_.msTotal = ms_since(t0);
    batchK_ = 0;
    finishStats();
    return st_;
}

template <class E, class G>
void Engine<E, G>::startStats()
{
    (void)deviceRounds();   // refreshes the host mirror
    ctl0_ = *h_ctl_;
}

template <class E, class G>
void Engine<E, G>::finishStats()
{
    SUBWAY_CHECK(hipStreamSynchronize(sx_));
    st_.innerRounds = deviceRounds() - ctl0_.rounds;
    const ScanCtl &c1 = *h_ctl_;
    st_.updates = c1.updates - ctl0_.updates;
    st_.workEdges = c1.work_sum - ctl0_.work_sum;
    st_.workEntries = c1.entry_sum - ctl0_.entry_sum;
    st_.workItems = c1.item_sum - ctl0_.item_sum;
    st_.laneWork = c1.lane_work - ctl0_.lane_work;
    st_.laneEntries = c1.lane_entries - ctl0_.lane_entries;
    st_.pairWorkEdges = c1.pair_work - ctl0_.pair_work;
    st_.pairWorkEntries = c1.pair_entries - ctl0_.pair_entries;
    st_.pairUpdates = c1.pair_updates - ctl0_.pair_updates;
    st_.blkWorkEdges = c1.blk_work - ctl0_.blk_work;
    st_.blkWorkEntries = c1.blk_entries - ctl0_.blk_entries;
    st_.blkUpdates = c1.blk_updates - ctl0_.blk_updates;
    st_.edgeBytes = (uint)eb_;
    if (pc_) {
        uint32_t mm = 0;
        SUBWAY_CHECK(hipMemcpy(&mm, &pc_->min_members, sizeof(mm), hipMemcpyDeviceToHost));
        st_.minMembers = mm == 0xFFFFFFFFu ? 0u : mm;
    }
    if (o_.profile) {
        ull c = 0;
        st_.msAdvance = pAdv_.total(st_.advanceLaunches);
        st_.msFrontier = pFr_.total(st_.frontierLaunches);
        st_.msTransfer = pXfer_.total(st_.transferOps);
        st_.msGenerate = pGen_.total(c);
        st_.msRoundsBlocked = pRoundsBlk_.total(st_.roundsLaunchesBlocked);
        st_.msRounds = pRounds_.total(st_.roundsLaunches) + st_.msRoundsBlocked;
        st_.roundsLaunches += st_.roundsLaunchesBlocked;
    }
}

template <class E, class G>
class SessionImpl : public Session {
public:
    SessionImpl(std::shared_ptr<HostCsr<E>> csr, const RunOptions &o) : eng_(std::move(csr), o), algo_(o.algo) {}
    RunStats run(uint source) override { return eng_.run(source); }
    void values(void *out) override
    {
        G &g = eng_.graph();
        const size_t esz = algo_ == PR ? sizeof(float) : sizeof(uint);
        const char *src = (const char *)g.value;
        parallel_for(g.num_nodes, [&](uint64_t b, uint64_t e, int) {
            std::memcpy((char *)out + b * esz, src + b * esz, (e - b) * esz);
        });
    }
    const void *valuesPtr() override { return eng_.graph().value; }
    ull numNodes() const override { return const_cast<SessionImpl *>(this)->eng_.graph().num_nodes; }
    RunStats runBatch(const uint *sources, int k) override { return eng_.runBatch(sources, k); }
    void setPolicy(const float *params) override { eng_.setPolicy(params); }
    void rows(uint *out) override
    {
        const uint32_t *r = eng_.batchRows();
        if (!r) throw Error(ERR_ARG, "no batched run yet");
        std::memcpy(out, r, (size_t)numNodes() * BATCH_LANES * sizeof(uint32_t));
    }
    int tuneBatch(const uint *sources, int k, double *timesMs) override
    {
        static const int cand[4] = {4, 8, 16, 32};
        int best = eng_.batchUnroll();
        double bestMs = 1e300;
        for (int i = 0; i < 4; i++) {
            eng_.setBatchUnroll(cand[i]);
            RunStats st = eng_.runBatch(sources, k);
            if (timesMs) timesMs[i] = st.msTotal;
            if (st.msTotal < bestMs) {
                bestMs = st.msTotal;
                best = cand[i];
            }
        }
        eng_.setBatchUnroll(best);
        return best;
    }

private:
    Engine<E, G> eng_;
    int algo_;
};

}  // namespace

template <class E>
std::unique_ptr<Session> make_session(std::shared_ptr<HostCsr<E>> csr, const RunOptions &o)
{
    if ((o.algo == SSSP || o.algo == SSWP) && sizeof(E) != 8)
        throw Error(ERR_ARG, "SSSP/SSWP need a weighted graph (.bwcsr/.wel)");
    if (o.algo == PR) return std::unique_ptr<Session>(new SessionImpl<E, GraphPR<E>>(std::move(csr), o));
    return std::unique_ptr<Session>(new SessionImpl<E, Graph<E>>(std::move(csr), o));
}

template <class E>
RunStats runLabel(Graph<E> &graph, const RunOptions &opt)
{
    if (opt.algo == PR) throw Error(ERR_ARG, "runLabel does not run PageRank");
    if ((opt.algo == SSSP || opt.algo == SSWP) && sizeof(E) != 8)
        throw Error(ERR_ARG, "SSSP/SSWP need a weighted graph (.bwcsr/.wel)");
    Engine<E, Graph<E>> eng(graph, opt);
    return eng.run(opt.source);
}

template <class E>
RunStats runPageRank(GraphPR<E> &graph, const RunOptions &opt)
{
    RunOptions o = opt;
    o.algo = PR;
    Engine<E, GraphPR<E>> eng(graph, o);
    return eng.run(0);
}

void release_shared_streams() { destroy_stream_pairs(); }

template std::unique_ptr<Session> make_session<OutEdge>(std::shared_ptr<HostCsr<OutEdge>>, const RunOptions &);
template std::unique_ptr<Session> make_session<OutEdgeWeighted>(std::shared_ptr<HostCsr<OutEdgeWeighted>>,
                                                                const RunOptions &);
template RunStats runLabel<OutEdge>(Graph<OutEdge> &, const RunOptions &);
template RunStats runLabel<OutEdgeWeighted>(Graph<OutEdgeWeighted> &, const RunOptions &);
template RunStats runPageRank<OutEdge>(GraphPR<OutEdge> &, const RunOptions &);
template RunStats runPageRank<OutEdgeWeighted>(GraphPR<OutEdgeWeighted> &, const RunOptions &);

}  // namespace subway
