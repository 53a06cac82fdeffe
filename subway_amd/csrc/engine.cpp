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
#include <cstring>
#include <iostream>
#include <memory>
#include <vector>

#include "host_util.hpp"
#include "kernels.hpp"
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
};

template <class E, class G>
class Engine {
public:
    Engine(G &g, const RunOptions &o) : g_(g), o_(o)
    {
        if (o_.numBuffers != 1 && o_.numBuffers != 2) throw Error(ERR_ARG, "numBuffers must be 1 or 2");
        if (o_.algo != PR && o_.source >= g_.num_nodes && o_.algo != CC)
            throw Error(ERR_ARG, "source vertex " + std::to_string(o_.source) + " out of range");
        if (o_.budgetBytes)
            sub_.reset(new Subgraph<E>(g_.num_nodes, g_.num_edges, o_.budgetBytes, o_.numBuffers));
        else
            sub_.reset(new Subgraph<E>(g_.num_nodes, g_.num_edges));
        gen_.reset(new SubgraphGenerator<E>(g_));
        part_.reset(new Partitioner<E>());
        SUBWAY_CHECK(hipStreamCreateWithFlags(&sc_, hipStreamNonBlocking));
        SUBWAY_CHECK(hipStreamCreateWithFlags(&sx_, hipStreamNonBlocking));
        gen_->stream = sc_;
        part_->stream = sc_;
        for (int b = 0; b < 2; b++) {
            SUBWAY_CHECK(hipEventCreateWithFlags(&evLoaded_[b], hipEventDisableTiming));
            SUBWAY_CHECK(hipEventCreateWithFlags(&evFree_[b], hipEventDisableTiming));
        }
        SUBWAY_CHECK(hipEventCreate(&evX0_));
        SUBWAY_CHECK(hipEventCreate(&evX1_));
        maxPart_ = sub_->max_partition_size;
        alloc_scan_scratch(fr_, ((ull)g_.num_nodes + FR_TILE - 1) / FR_TILE + 1);
        ull maxEntries = std::min<ull>((ull)g_.num_nodes, maxPart_) + 1;
        SUBWAY_CHECK(hipMalloc(&entries_, maxEntries * sizeof(Entry)));
        ull maxWork = std::max<ull>(maxPart_, (ull)g_.csr->maxDegree);
        SUBWAY_CHECK(hipMalloc(&chunkMap_, (maxWork / ADV_CHUNK + 2) * sizeof(uint32_t)));
        SUBWAY_CHECK(hipHostMalloc((void **)&h_ctl_, sizeof(ScanCtl), 0));
        int dev = 0;
        SUBWAY_CHECK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        SUBWAY_CHECK(hipGetDeviceProperties(&prop, dev));
        advGrid_ = o_.advanceGrid > 0 ? o_.advanceGrid : prop.multiProcessorCount * 4;
    }

    ~Engine()
    {
        if (sc_) (void)hipStreamSynchronize(sc_);
        if (sx_) (void)hipStreamSynchronize(sx_);
        free_scan_scratch(fr_);
        if (entries_) (void)hipFree(entries_);
        if (chunkMap_) (void)hipFree(chunkMap_);
        if (h_ctl_) (void)hipHostFree(h_ctl_);
        for (int b = 0; b < 2; b++) {
            (void)hipEventDestroy(evLoaded_[b]);
            (void)hipEventDestroy(evFree_[b]);
        }
        (void)hipEventDestroy(evX0_);
        (void)hipEventDestroy(evX1_);
        part_.reset();
        gen_.reset();
        sub_.reset();
        if (sc_) (void)hipStreamDestroy(sc_);
        if (sx_) (void)hipStreamDestroy(sx_);
    }

    RunStats run();

private:
    G &g_;
    RunOptions o_;
    std::unique_ptr<Subgraph<E>> sub_;
    std::unique_ptr<SubgraphGenerator<E>> gen_;
    std::unique_ptr<Partitioner<E>> part_;
    hipStream_t sc_ = nullptr, sx_ = nullptr;
    hipEvent_t evLoaded_[2], evFree_[2], evX0_, evX1_;
    ScanScratch fr_;
    Entry *entries_ = nullptr;
    uint32_t *chunkMap_ = nullptr;
    ScanCtl *h_ctl_ = nullptr;
    ull maxPart_ = 0;
    int advGrid_ = 0;
    RunStats st_;
    std::vector<Unit> units_;
    EventPairs pAdv_, pFr_, pXfer_, pGen_;

    static constexpr bool isPR = std::is_same<G, GraphPR<E>>::value;

    void initState();
    void generateAndPartition();
    void buildUnits();
    void issue(size_t u);
    void compute(size_t u);
    void frontier(int first, uint8_t *labA, uint8_t *labB, const Unit &un);
    void advance(uint8_t *labOut, ull w0, ull w1, ull shift, int buf);
    void asyncRounds(const Unit &un, int buf);
    ull deviceRounds();
};

template <class E, class G>
void Engine<E, G>::initState()
{
    const uint n = g_.num_nodes;
    if constexpr (isPR) {
        for (uint i = 0; i < n; i++) {
            g_.delta[i] = o_.initPR;
            g_.value[i] = 0.0f;
        }
        SUBWAY_CHECK(hipMemcpy(g_.d_value, g_.value, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
        SUBWAY_CHECK(hipMemcpy(g_.d_delta, g_.delta, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    } else {
        const bool async = o_.mode == ASYNC;
        for (uint i = 0; i < n; i++) {
            if (o_.algo == CC) g_.value[i] = i;
            else if (o_.algo == SSWP) g_.value[i] = 0;
            else g_.value[i] = DIST_INFINITY;
            g_.label1[i] = async;
            g_.label2[i] = (!async && o_.algo == CC);
        }
        if (o_.algo != CC) {
            g_.value[o_.source] = (o_.algo == SSWP) ? DIST_INFINITY : 0;
            if (!async) g_.label2[o_.source] = true;
        }
        SUBWAY_CHECK(hipMemcpy(g_.d_value, g_.value, (size_t)n * sizeof(uint), hipMemcpyHostToDevice));
        SUBWAY_CHECK(hipMemcpy(g_.d_label1, g_.label1, (size_t)n, hipMemcpyHostToDevice));
        SUBWAY_CHECK(hipMemcpy(g_.d_label2, g_.label2, (size_t)n, hipMemcpyHostToDevice));
    }
    SUBWAY_CHECK(hipDeviceSynchronize());
}

template <class E, class G>
void Engine<E, G>::generateAndPartition()
{
    if (o_.profile) pGen_.begin(sc_);
    if constexpr (isPR) gen_->enqueue(g_, *sub_, o_.acc);
    else gen_->enqueue(g_, *sub_);
    if (o_.profile) pGen_.end(sc_);
    part_->partition(*sub_, gen_->deviceTotals(), maxPart_);
    sub_->numActiveNodes = (uint)part_->numActiveNodes;
    sub_->numActiveEdges = part_->numActiveEdges;
}

template <class E, class G>
void Engine<E, G>::buildUnits()
{
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
            units_.push_back(u);
            continue;
        }
        if (cnt == 1 && edges > maxPart_) {
            const uint v = p.firstVertex[i];
            const ull base = g_.nodePointer[v];
            for (ull c = 0; c * maxPart_ < edges; c++) {
                Unit h{};
                h.part = i;
                h.hub = true;
                h.chunk = c;
                h.dma = true;
                h.hostEdge = base + c * maxPart_;
                h.count = std::min(maxPart_, edges - c * maxPart_);
                units_.push_back(h);
            }
            continue;
        }
        const bool contiguous = (ull)p.lastVertex[i] - (ull)p.firstVertex[i] + 1 == cnt;
        if (contiguous && o_.transfer != TRANSFER_ZEROCOPY) {
            u.dma = true;
            u.hostEdge = g_.nodePointer[p.firstVertex[i]];
        } else {
            u.dma = false;
        }
        u.count = edges;
        u.kbeg = p.fromNode[i];
        u.kend = p.fromNode[i] + cnt;
        u.edgeBase = p.fromEdge[i];
        units_.push_back(u);
    }
}

template <class E, class G>
void Engine<E, G>::issue(size_t idx)
{
    const Unit &u = units_[idx];
    const int b = (int)(idx % (size_t)sub_->numBuffers);
    SUBWAY_CHECK(hipStreamWaitEvent(sx_, evFree_[b], 0));
    if (u.count) {
        if (o_.profile) pXfer_.begin(sx_);
        E *dst = sub_->d_buffers[b];
        if (u.dma) {
            SUBWAY_CHECK(hipMemcpyAsync(dst, g_.edgeList + u.hostEdge, (size_t)u.count * sizeof(E),
                                        hipMemcpyHostToDevice, sx_));
            st_.h2dDmaBytes += u.count * sizeof(E);
        } else {
            launch_gather((int)sizeof(E), sub_->d_activeNodes, sub_->d_activeNodesPointer, u.kbeg, u.kend,
                          g_.d_nodePointer, g_.d_edgeList, dst, u.edgeBase, u.count, sx_);
            st_.h2dGatherBytes += u.count * sizeof(E);
        }
        st_.h2dBytes += u.count * sizeof(E);
        if (o_.profile) pXfer_.end(sx_);
    }
    SUBWAY_CHECK(hipEventRecord(evLoaded_[b], sx_));
}

template <class E, class G>
void Engine<E, G>::frontier(int first, uint8_t *labA, uint8_t *labB, const Unit &un)
{
    const Partitioner<E> &p = *part_;
    FrontierArgs a{};
    a.algo = o_.algo;
    a.n = p.partitionNodeSize[un.part];
    a.nodes = sub_->d_activeNodes + p.fromNode[un.part];
    a.ptr = sub_->d_activeNodesPointer + p.fromNode[un.part];
    a.fromEdge = p.fromEdge[un.part];
    a.labA = labA;
    a.labB = labB;
    if constexpr (isPR) {
        a.value = (uint32_t *)g_.d_value;
        a.delta = g_.d_delta;
        a.acc = o_.acc;
    } else {
        a.value = g_.d_value;
    }
    a.entries = entries_;
    a.chunkMap = chunkMap_;
    a.first = first;
    if (o_.profile) pFr_.begin(sc_);
    launch_frontier(a, fr_, sc_);
    if (o_.profile) pFr_.end(sc_);
}

template <class E, class G>
void Engine<E, G>::advance(uint8_t *labOut, ull w0, ull w1, ull shift, int buf)
{
    AdvanceArgs a{};
    a.algo = o_.algo;
    a.edgeBytes = (int)sizeof(E);
    a.entries = entries_;
    a.chunkMap = chunkMap_;
    a.edges = sub_->d_buffers[buf];
    if constexpr (isPR) {
        a.delta = g_.d_delta;
    } else {
        a.value = g_.d_value;
        a.labelOut = labOut;
    }
    a.w0 = w0;
    a.w1 = w1;
    a.shift = shift;
    a.grid = advGrid_;
    if (o_.profile) pAdv_.begin(sc_);
    launch_advance(a, fr_.ctl, sc_);
    if (o_.profile) pAdv_.end(sc_);
}

template <class E, class G>
void Engine<E, G>::asyncRounds(const Unit &un, int buf)
{
    uint8_t *l1 = nullptr, *l2 = nullptr;
    if constexpr (!isPR) {
        l1 = (uint8_t *)g_.d_label1;
        l2 = (uint8_t *)g_.d_label2;
    }
    ull r = 0;
    int batch = 2;
    for (;;) {
        for (int q = 0; q < batch; q++) {
            r++;
            const bool odd = (r & 1) != 0;
            if (r == 1) frontier(1, l1, l2, un);
            else frontier(0, odd ? l1 : l2, nullptr, un);
            advance(odd ? l2 : l1, 0, ~0ull, 0, buf);
        }
        SUBWAY_CHECK(hipMemcpyAsync(&h_ctl_->stop, &fr_.ctl->stop, sizeof(uint32_t), hipMemcpyDeviceToHost, sc_));
        SUBWAY_CHECK(hipStreamSynchronize(sc_));
        if (h_ctl_->stop) break;
        batch = std::min(batch * 2, std::max(1, o_.maxRoundBatch));
    }
}

template <class E, class G>
void Engine<E, G>::compute(size_t idx)
{
    const Unit &u = units_[idx];
    const int b = (int)(idx % (size_t)sub_->numBuffers);
    SUBWAY_CHECK(hipStreamWaitEvent(sc_, evLoaded_[b], 0));
    uint8_t *l1 = nullptr, *l2 = nullptr;
    if constexpr (!isPR) {
        l1 = (uint8_t *)g_.d_label1;
        l2 = (uint8_t *)g_.d_label2;
    }
    const bool async = o_.mode == ASYNC;
    if (u.hub) {
        // one round over the oversized vertex, streamed chunk by chunk: the
        // frontier kernel runs once (clearing its label, fixing its value),
        // then each chunk is relaxed through a window of the work space
        if (u.chunk == 0) {
            if (isPR) frontier(1, nullptr, nullptr, u);
            else if (async) frontier(1, l1, l2, u);
            else frontier(1, l2, nullptr, u);
        }
        const ull w0 = u.chunk * maxPart_;
        advance(l2, w0, w0 + u.count, w0, b);
    } else if (!async) {
        if (isPR) frontier(1, nullptr, nullptr, u);
        else frontier(1, l2, nullptr, u);
        advance(l2, 0, ~0ull, 0, b);
    } else {
        asyncRounds(u, b);
    }
    SUBWAY_CHECK(hipEventRecord(evFree_[b], sc_));
    if (o_.verbose && (!u.hub || u.chunk == 0)) {
        std::cout << "Global Iteration " << st_.globalIterations << ", Partition " << u.part << " of "
                  << part_->numPartitions << ": " << part_->partitionNodeSize[u.part] << " vertices, "
                  << part_->partitionEdgeSize[u.part] << " edges\n";
    }
}

template <class E, class G>
ull Engine<E, G>::deviceRounds()
{
    SUBWAY_CHECK(hipMemcpyAsync(h_ctl_, fr_.ctl, sizeof(ScanCtl), hipMemcpyDeviceToHost, sc_));
    SUBWAY_CHECK(hipStreamSynchronize(sc_));
    return h_ctl_->rounds;
}

template <class E, class G>
RunStats Engine<E, G>::run()
{
    initState();
    const uint n = g_.num_nodes;
    ull rounds0 = deviceRounds(), updates0 = h_ctl_->updates;
    ull work0 = h_ctl_->work_sum, entries0 = h_ctl_->entry_sum;
    for (int b = 0; b < 2; b++) SUBWAY_CHECK(hipEventRecord(evFree_[b], sc_));
    SUBWAY_CHECK(hipDeviceSynchronize());

    const auto t0 = clk::now();
    generateAndPartition();
    if constexpr (!isPR) {
        if (o_.mode == ASYNC && o_.algo != CC) {
            // every vertex was active for the first generate; from now on only
            // the source is labelled (sssp-async.cu:59-64)
            launch_fill_u8((uint8_t *)g_.d_label1, 0, n, sc_);
            static const uint8_t one = 1;
            SUBWAY_CHECK(hipMemcpyAsync((uint8_t *)g_.d_label1 + o_.source, &one, 1, hipMemcpyHostToDevice, sc_));
        }
    }
    while (sub_->numActiveNodes > 0) {
        const auto ti = clk::now();
        st_.globalIterations++;
        IterRecord rec{};
        rec.iteration = st_.globalIterations;
        rec.activeNodes = sub_->numActiveNodes;
        rec.activeEdges = sub_->numActiveEdges;
        rec.partitions = part_->numPartitions;
        const ull h2d0 = st_.h2dBytes;
        buildUnits();
        const size_t nu = units_.size();
        const size_t ahead = (size_t)sub_->numBuffers - 1;
        size_t issued = 0;
        for (size_t u = 0; u < nu; u++) {
            while (issued < nu && issued <= u + ahead) issue(issued++);
            compute(u);
        }
        st_.partitions += part_->numPartitions;
        rec.h2dBytes = st_.h2dBytes - h2d0;
        generateAndPartition();
        rec.msIteration = ms_since(ti);
        st_.trace.push_back(rec);
    }
    if constexpr (isPR)
        SUBWAY_CHECK(hipMemcpyAsync(g_.value, g_.d_value, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, sc_));
    else
        SUBWAY_CHECK(hipMemcpyAsync(g_.value, g_.d_value, (size_t)n * sizeof(uint), hipMemcpyDeviceToHost, sc_));
    SUBWAY_CHECK(hipStreamSynchronize(sc_));
    st_.msTotal = ms_since(t0);
    SUBWAY_CHECK(hipStreamSynchronize(sx_));
    st_.innerRounds = deviceRounds() - rounds0;
    st_.updates = h_ctl_->updates - updates0;
    st_.workEdges = h_ctl_->work_sum - work0;
    st_.workEntries = h_ctl_->entry_sum - entries0;
    if (o_.profile) {
        ull c = 0;
        st_.msAdvance = pAdv_.total(st_.advanceLaunches);
        st_.msFrontier = pFr_.total(st_.frontierLaunches);
        st_.msTransfer = pXfer_.total(st_.transferOps);
        st_.msGenerate = pGen_.total(c);
    }
    return st_;
}

}  // namespace

template <class E>
RunStats runLabel(Graph<E> &graph, const RunOptions &opt)
{
    if (opt.algo == PR) throw Error(ERR_ARG, "runLabel does not run PageRank");
    if ((opt.algo == SSSP || opt.algo == SSWP) && sizeof(E) != 8)
        throw Error(ERR_ARG, "SSSP/SSWP need a weighted graph (.bwcsr/.wel)");
    Engine<E, Graph<E>> eng(graph, opt);
    return eng.run();
}

template <class E>
RunStats runPageRank(GraphPR<E> &graph, const RunOptions &opt)
{
    RunOptions o = opt;
    o.algo = PR;
    Engine<E, GraphPR<E>> eng(graph, o);
    return eng.run();
}

template RunStats runLabel<OutEdge>(Graph<OutEdge> &, const RunOptions &);
template RunStats runLabel<OutEdgeWeighted>(Graph<OutEdgeWeighted> &, const RunOptions &);
template RunStats runPageRank<OutEdge>(GraphPR<OutEdge> &, const RunOptions &);
template RunStats runPageRank<OutEdgeWeighted>(GraphPR<OutEdgeWeighted> &, const RunOptions &);

}  // namespace subway
