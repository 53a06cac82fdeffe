// capi.cpp -- extern "C" facade over the C++ classes (see subway_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "host_util.hpp"
#include "kernels.hpp"
#include "mlp.hpp"
#include "subway/engine.hpp"
#include "subway/gpu_error_check.hpp"
#include "subway/graph.hpp"
#include "subway_amd.h"

namespace subway {
template <class E>
std::shared_ptr<HostCsr<E>> rmat_csr(int scale, int edgeFactor, double a, double b, double c, uint64_t seed,
                                     bool symmetrize, uint32_t wmin, uint32_t wmax, bool splitLayout);
template <class E>
std::shared_ptr<HostCsr<E>> rmat_csr_gpu(int scale, int edgeFactor, double a, double b, double c, uint64_t seed,
                                         bool symmetrize, uint32_t wmin, uint32_t wmax, bool splitLayout);
}

using namespace subway;

struct subway_graph {
    bool weighted = false;
    std::shared_ptr<HostCsr<OutEdge>> u;
    std::shared_ptr<HostCsr<OutEdgeWeighted>> w;
    ull n() const { return weighted ? w->n : u->n; }
    ull m() const { return weighted ? w->m : u->m; }
};

static thread_local std::string g_last_error;

namespace subway {
// the calling thread's message for subway_last_error (entry points defined
// in other translation units report through it)
void set_last_error(const char *msg) { g_last_error = msg ? msg : ""; }
}  // namespace subway

template <class F>
static int32_t guard(F &&f)
{
    try {
        f();
        return SUBWAY_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code ? e.code : SUBWAY_ERR_INTERNAL;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return SUBWAY_ERR_NOMEM;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return SUBWAY_ERR_INTERNAL;
    }
}

template <class T>
struct DevBuf {
    T *p = nullptr;
    explicit DevBuf(size_t n)
    {
        SUBWAY_CHECK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
    }
    ~DevBuf() { if (p) (void)hipFree(p); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

extern "C" {

int32_t subway_version(void) { return 100; }

const char *subway_last_error(void) { return g_last_error.c_str(); }

void subway_default_options(subway_options *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->algo = SUBWAY_BFS;
    o->mode = SUBWAY_SYNC;
    o->num_buffers = 2;
    o->acc = 0.01f;
    o->init_pr = 0.15f;
    o->theta = 0.5f;
}

void subway_struct_sizes(uint32_t *options, uint32_t *stats, uint32_t *iter_record)
{
    if (options) *options = (uint32_t)sizeof(subway_options);
    if (stats) *stats = (uint32_t)sizeof(subway_stats);
    if (iter_record) *iter_record = (uint32_t)sizeof(subway_iter_record);
}

void subway_shutdown(void) { release_shared_streams(); }

int32_t subway_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int32_t subway_load(const char *path, int32_t weighted, subway_graph **out)
{
    return guard([&] {
        if (!path || !out) throw Error(ERR_ARG, "null argument");
        std::unique_ptr<subway_graph> g(new subway_graph);
        g->weighted = weighted != 0;
        if (g->weighted) g->w = read_csr_file<OutEdgeWeighted>(path, weighted == 2);
        else g->u = read_csr_file<OutEdge>(path);
        *out = g.release();
    });
}

}  // extern "C"

template <class E>
static std::shared_ptr<HostCsr<E>> csr_from_arrays(uint64_t n, uint64_t m, const uint64_t *ptr, const uint32_t *ends,
                                                   const uint32_t *weights)
{
    auto c = std::make_shared<HostCsr<E>>();
    c->allocate(n, m);
    if (ptr[0] != 0 || ptr[n] != m) throw Error(ERR_FORMAT, "node_pointer must start at 0 and end at m");
    for (uint64_t i = 0; i <= n; i++) c->ptr[i] = ptr[i];
    for (uint64_t e = 0; e < m; e++) {
        if (ends[e] >= n) throw Error(ERR_FORMAT, "edge endpoint out of range");
        c->edges[e].end = ends[e];
        if constexpr (sizeof(E) == 8) reinterpret_cast<uint32_t *>(&c->edges[e])[1] = weights[e];
    }
    c->finalize();
    return c;
}

extern "C" {

int32_t subway_graph_from_csr(uint64_t n, uint64_t m, const uint64_t *node_pointer, const uint32_t *ends,
                              const uint32_t *weights, subway_graph **out)
{
    return guard([&] {
        if (!node_pointer || !out || (m && !ends)) throw Error(ERR_ARG, "null argument");
        std::unique_ptr<subway_graph> g(new subway_graph);
        g->weighted = weights != nullptr;
        if (g->weighted) g->w = csr_from_arrays<OutEdgeWeighted>(n, m, node_pointer, ends, weights);
        else g->u = csr_from_arrays<OutEdge>(n, m, node_pointer, ends, nullptr);
        *out = g.release();
    });
}

int32_t subway_graph_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                          int32_t weighted, int32_t symmetrize, uint32_t wmin, uint32_t wmax, subway_graph **out)
{
    return guard([&] {
        if (!out) throw Error(ERR_ARG, "null argument");
        std::unique_ptr<subway_graph> g(new subway_graph);
        g->weighted = weighted != 0;
        if (g->weighted)
            g->w = rmat_csr<OutEdgeWeighted>(scale, edge_factor, a, b, c, seed, symmetrize != 0, wmin, wmax,
                                             weighted == 2);
        else
            g->u = rmat_csr<OutEdge>(scale, edge_factor, a, b, c, seed, symmetrize != 0, wmin, wmax, false);
        *out = g.release();
    });
}

int32_t subway_graph_rmat_device(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                                 int32_t weighted, int32_t symmetrize, uint32_t wmin, uint32_t wmax,
                                 subway_graph **out)
{
    return guard([&] {
        if (!out) throw Error(ERR_ARG, "null argument");
        std::unique_ptr<subway_graph> g(new subway_graph);
        g->weighted = weighted != 0;
        if (g->weighted)
            g->w = rmat_csr_gpu<OutEdgeWeighted>(scale, edge_factor, a, b, c, seed, symmetrize != 0, wmin, wmax,
                                                 weighted == 2);
        else
            g->u = rmat_csr_gpu<OutEdge>(scale, edge_factor, a, b, c, seed, symmetrize != 0, wmin, wmax, false);
        *out = g.release();
    });
}

int32_t subway_graph_reorder(subway_graph *g, int32_t order)
{
    return guard([&] {
        if (!g) throw Error(ERR_ARG, "null graph");
        if (order != SUBWAY_ORDER_IN_DEGREE) throw Error(ERR_ARG, "unknown vertex order");
        // sessions keep their own references to the CSR (and device state built from it)
        const long users = g->weighted ? g->w.use_count() : g->u.use_count();
        if (users > 1) throw Error(ERR_ARG, "renumber a graph before any session is created on it");
        if (g->weighted) reorder_by_in_degree(*g->w);
        else reorder_by_in_degree(*g->u);
    });
}

int32_t subway_graph_build_blocks(subway_graph *g, int32_t flags)
{
    return guard([&] {
        if (!g) throw Error(ERR_ARG, "null graph");
        if (flags & ~SUBWAY_BLOCKS_UNPIN_SPLIT) throw Error(ERR_ARG, "unknown block layout flags");
        const long users = g->weighted ? g->w.use_count() : g->u.use_count();
        if (users > 1) throw Error(ERR_ARG, "build the block layout before any session is created on the graph");
        const bool unpin = (flags & SUBWAY_BLOCKS_UNPIN_SPLIT) != 0;
        if (g->weighted) build_blocks(*g->w, unpin);
        else build_blocks(*g->u, unpin);
    });
}

int32_t subway_graph_blocks_view(const subway_graph *g, const uint64_t **row_ptr, const uint32_t **words,
                                 uint64_t *num_words, uint32_t *weight_code)
{
    return guard([&] {
        if (!g) throw Error(ERR_ARG, "null graph");
        const uint32_t *w = g->weighted ? g->w->blk : g->u->blk;
        if (!w) throw Error(ERR_ARG, "the graph holds no block layout");
        if (row_ptr) *row_ptr = reinterpret_cast<const uint64_t *>(g->weighted ? g->w->blkPtr : g->u->blkPtr);
        if (words) *words = w;
        if (num_words) *num_words = g->weighted ? g->w->blkWords : g->u->blkWords;
        if (weight_code) *weight_code = g->weighted ? g->w->blkCode : g->u->blkCode;
    });
}

int32_t subway_graph_split_tail(const subway_graph *g, uint64_t *tail_edge)
{
    return guard([&] {
        if (!g || !tail_edge) throw Error(ERR_ARG, "null argument");
        if (g->weighted) *tail_edge = g->w->tailDst ? g->w->tailEdge : g->w->m;
        else *tail_edge = g->u->tailDst ? g->u->tailEdge : g->u->m;
    });
}

int32_t subway_graph_order(const subway_graph *g, const uint32_t **perm, const uint32_t **iperm)
{
    return guard([&] {
        if (!g) throw Error(ERR_ARG, "null graph");
        const uint32_t *p = g->weighted ? g->w->perm : g->u->perm;
        const uint32_t *ip = g->weighted ? g->w->iperm : g->u->iperm;
        if (perm) *perm = p;
        if (iperm) *iperm = ip;
    });
}

int32_t subway_graph_split_view(const subway_graph *g, const uint64_t **node_pointer, const uint32_t **ends,
                                const uint8_t **weights)
{
    return guard([&] {
        if (!g) throw Error(ERR_ARG, "null graph");
        if (!g->weighted || !g->w->splitDst) throw Error(ERR_ARG, "the graph holds no split edge arrays");
        if (node_pointer) *node_pointer = reinterpret_cast<const uint64_t *>(g->w->ptr);
        if (ends) *ends = g->w->splitDst;
        if (weights) *weights = g->w->splitW;
    });
}

int32_t subway_graph_save(const subway_graph *g, const char *path)
{
    return guard([&] {
        if (!g || !path) throw Error(ERR_ARG, "null argument");
        if (g->weighted) write_csr_file(*g->w, path);
        else write_csr_file(*g->u, path);
    });
}

int32_t subway_graph_info(const subway_graph *g, uint64_t *n, uint64_t *m, int32_t *weighted, uint32_t *max_degree,
                          uint32_t *max_degree_vertex)
{
    return guard([&] {
        if (!g) throw Error(ERR_ARG, "null graph");
        if (n) *n = g->n();
        if (m) *m = g->m();
        if (weighted) *weighted = g->weighted;
        const uint *deg = g->weighted ? g->w->degree : g->u->degree;
        uint md = g->weighted ? g->w->maxDegree : g->u->maxDegree;
        if (max_degree) *max_degree = md;
        if (max_degree_vertex) {
            // the smallest original id among the vertices of largest out-degree
            const uint32_t *ip = g->weighted ? g->w->iperm : g->u->iperm;
            uint64_t at = ~0ull;
            for (uint64_t i = 0; i < g->n(); i++)
                if (deg[i] == md) {
                    const uint64_t v = ip ? ip[i] : i;
                    if (v < at) at = v;
                    if (!ip) break;
                }
            *max_degree_vertex = (uint32_t)at;
        }
    });
}

int32_t subway_graph_export(const subway_graph *g, uint64_t *node_pointer, uint32_t *ends, uint32_t *weights)
{
    return guard([&] {
        if (!g) throw Error(ERR_ARG, "null graph");
        const ull n = g->n(), m = g->m();
        const ull *ptr = g->weighted ? g->w->ptr : g->u->ptr;
        if (node_pointer) std::memcpy(node_pointer, ptr, (n + 1) * sizeof(ull));
        if (!ends && !weights) return;
        parallel_for(m, [&](uint64_t b, uint64_t e1, int) {
            for (ull e = b; e < e1; e++) {
                if (g->weighted) {
                    const OutEdgeWeighted x = g->w->edgeAt(e);
                    if (ends) ends[e] = x.end;
                    if (weights) weights[e] = x.w8;
                } else {
                    if (ends) ends[e] = g->u->edges[e].end;
                    if (weights) weights[e] = 1;
                }
            }
        });
    });
}

void subway_graph_free(subway_graph *g) { delete g; }

static RunOptions to_options(const subway_options *o)
{
    RunOptions r;
    r.algo = o->algo;
    r.mode = o->mode;
    r.source = o->source;
    r.budgetBytes = o->budget_bytes;
    r.numBuffers = o->num_buffers ? o->num_buffers : 2;
    r.acc = o->acc;
    r.initPR = o->init_pr;
    r.transfer = o->transfer;
    r.policy = o->policy;
    r.verbose = o->verbose;
    r.advanceGrid = o->advance_grid;
    r.profile = o->profile;
    r.persistent = o->rounds_mode == 0;
    r.batch = o->batch;
    r.batchUnroll = o->batch_unroll;
    r.theta = o->theta;
    r.roundsGrid = o->rounds_grid;
    if (o->max_rounds > 0) r.maxRounds = o->max_rounds;
    r.spinLimit = o->spin_limit;
    r.joinWindow = o->join_window;
    r.unpackedEdges = o->unpacked_edges;
    r.noResident = o->no_resident;
    r.residentMinBytes = o->resident_min_bytes;
    r.pushOnly = o->push_only;
    r.noBlocks = o->no_blocks;
    r.earlyReadback = o->early_readback;
    if (r.algo < BFS || r.algo > PR) throw Error(ERR_ARG, "unknown algorithm");
    if (r.mode != SYNC && r.mode != ASYNC) throw Error(ERR_ARG, "unknown mode");
    return r;
}

}  // extern "C"

struct subway_session {
    std::shared_ptr<void> keep;          // the graph's CSR stays alive with the session
    std::unique_ptr<Session> impl;
};

static void fill_stats(const RunStats &s, subway_stats *stats, subway_iter_record *trace, uint32_t trace_cap)
{
    uint32_t k = 0;
    if (trace) {
        for (const IterRecord &r : s.trace) {
            if (k >= trace_cap) break;
            trace[k].iteration = r.iteration;
            trace[k].partitions = r.partitions;
            trace[k].active_nodes = r.activeNodes;
            trace[k].active_edges = r.activeEdges;
            trace[k].h2d_bytes = r.h2dBytes;
            trace[k].ms_iteration = r.msIteration;
            std::memcpy(trace[k].features, r.features, sizeof(trace[k].features));
            std::memcpy(trace[k].heads, r.heads, sizeof(trace[k].heads));
            trace[k].theta = r.theta;
            trace[k].rounds_cap = r.roundsCap;
            trace[k].on_plain = r.onPlain;
            k++;
        }
    }
    if (!stats) return;
    std::memset(stats, 0, sizeof(*stats));
    stats->ms_total = s.msTotal;
    stats->ms_transfer = s.msTransfer;
    stats->ms_advance = s.msAdvance;
    stats->ms_frontier = s.msFrontier;
    stats->ms_generate = s.msGenerate;
    stats->ms_rounds = s.msRounds;
    stats->rounds_launches = s.roundsLaunches;
    stats->h2d_bytes = s.h2dBytes;
    stats->h2d_dma_bytes = s.h2dDmaBytes;
    stats->h2d_gather_bytes = s.h2dGatherBytes;
    stats->inner_rounds = s.innerRounds;
    stats->partitions = s.partitions;
    stats->updates = s.updates;
    stats->advance_launches = s.advanceLaunches;
    stats->frontier_launches = s.frontierLaunches;
    stats->transfer_ops = s.transferOps;
    stats->work_edges = s.workEdges;
    stats->work_entries = s.workEntries;
    stats->global_iterations = s.globalIterations;
    stats->trace_len = k;
    stats->work_items = s.workItems;
    stats->mispredictions = s.mispredictions;
    stats->fallback_iteration = s.fallbackIteration;
    stats->edge_bytes = s.edgeBytes;
    stats->min_members = s.minMembers;
    stats->lane_work = s.laneWork;
    stats->lane_entries = s.laneEntries;
    stats->pair_work_edges = s.pairWorkEdges;
    stats->pair_work_entries = s.pairWorkEntries;
    stats->pair_updates = s.pairUpdates;
    stats->resident_visits = s.residentVisits;
    stats->resident_overlaps = s.residentOverlaps;
    stats->pull_rounds = s.pullRounds;
    stats->pull_found = s.pullFound;
    stats->blocked_units = s.blockedUnits;
    stats->h2d_blocked_bytes = s.h2dBlockedBytes;
    stats->blocked_edges = s.blockedEdges;
    stats->ms_rounds_blocked = s.msRoundsBlocked;
    stats->rounds_launches_blocked = s.roundsLaunchesBlocked;
    stats->blk_work_edges = s.blkWorkEdges;
    stats->blk_work_entries = s.blkWorkEntries;
    stats->blk_updates = s.blkUpdates;
    stats->early_readback = s.earlyReadback;
    stats->values_patched = s.valuesPatched;
}

extern "C" {

int32_t subway_session_create(const subway_graph *g, const subway_options *opt, subway_session **out)
{
    return guard([&] {
        if (!g || !opt || !out) throw Error(ERR_ARG, "null argument");
        RunOptions o = to_options(opt);
        if ((o.algo == SSSP || o.algo == SSWP) && !g->weighted)
            throw Error(ERR_ARG, "SSSP/SSWP need a weighted graph");
        std::unique_ptr<subway_session> s(new subway_session);
        if (g->weighted) {
            s->keep = g->w;
            s->impl = make_session(g->w, o);
        } else {
            s->keep = g->u;
            s->impl = make_session(g->u, o);
        }
        *out = s.release();
    });
}

int32_t subway_session_run(subway_session *s, uint32_t source, void *values, subway_stats *stats,
                           subway_iter_record *trace, uint32_t trace_cap)
{
    return guard([&] {
        if (!s) throw Error(ERR_ARG, "null session");
        RunStats r = s->impl->run(source);
        if (values) s->impl->values(values);
        fill_stats(r, stats, trace, trace_cap);
    });
}

const void *subway_session_values(const subway_session *s) { return s ? s->impl->valuesPtr() : nullptr; }

int32_t subway_session_set_policy(subway_session *s, const float *params)
{
    return guard([&] {
        if (!s) throw Error(ERR_ARG, "null session");
        s->impl->setPolicy(params);
    });
}

void subway_session_free(subway_session *s) { delete s; }

int32_t subway_session_run_batch(subway_session *s, const uint32_t *sources, int32_t k, uint32_t *rows,
                                 subway_stats *stats, subway_iter_record *trace, uint32_t trace_cap)
{
    return guard([&] {
        if (!s || !sources) throw Error(ERR_ARG, "null argument");
        RunStats r = s->impl->runBatch(sources, k);
        if (rows) s->impl->rows(rows);
        fill_stats(r, stats, trace, trace_cap);
    });
}

int32_t subway_session_tune_batch(subway_session *s, const uint32_t *sources, int32_t k, int32_t *best_unroll,
                                  double *times_ms)
{
    return guard([&] {
        if (!s || !sources) throw Error(ERR_ARG, "null argument");
        int b = s->impl->tuneBatch(sources, k, times_ms);
        if (best_unroll) *best_unroll = b;
    });
}

int32_t subway_run(const subway_graph *g, const subway_options *opt, void *values, subway_stats *stats,
                   subway_iter_record *trace, uint32_t trace_cap)
{
    subway_session *s = nullptr;
    int32_t rc = subway_session_create(g, opt, &s);
    if (rc) return rc;
    rc = subway_session_run(s, opt->source, values, stats, trace, trace_cap);
    subway_session_free(s);
    return rc;
}

int32_t subway_generate(uint64_t n, const uint8_t *label1, const uint8_t *label2, const uint32_t *out_degree,
                        uint32_t *active_nodes, uint64_t *active_nodes_pointer, uint64_t *num_active)
{
    return guard([&] {
        if (!n || !label1 || !label2 || !out_degree || !active_nodes || !active_nodes_pointer)
            throw Error(ERR_ARG, "null argument or empty graph");
        DevBuf<uint8_t> l1(n), l2(n);
        DevBuf<uint32_t> deg(n), an(n);
        DevBuf<ull> ap(n + 1);
        SUBWAY_CHECK(hipMemcpy(l1.p, label1, n, hipMemcpyHostToDevice));
        SUBWAY_CHECK(hipMemcpy(l2.p, label2, n, hipMemcpyHostToDevice));
        SUBWAY_CHECK(hipMemcpy(deg.p, out_degree, n * 4, hipMemcpyHostToDevice));
        ScanScratch s;
        alloc_scan_scratch(s, (n + GEN_TILE - 1) / GEN_TILE);
        struct F { ScanScratch &s; ~F() { free_scan_scratch(s); } } f{s};
        launch_generate_labels(n, l1.p, l2.p, deg.p, an.p, ap.p, s, nullptr);
        ScanCtl ctl;
        SUBWAY_CHECK(hipMemcpy(&ctl, s.ctl, sizeof(ctl), hipMemcpyDeviceToHost));
        if (ctl.error)
            throw Error(ERR_INTERNAL, std::string("generator: device wait gave up (") + wait_error_name(ctl.error) +
                                          ")");
        SUBWAY_CHECK(hipMemcpy(active_nodes, an.p, ctl.count * 4, hipMemcpyDeviceToHost));
        SUBWAY_CHECK(hipMemcpy(active_nodes_pointer, ap.p, (ctl.count + 1) * 8, hipMemcpyDeviceToHost));
        if (num_active) *num_active = ctl.count;
    });
}

int32_t subway_generate_pr(uint64_t n, const float *delta, float acc, const uint32_t *out_degree,
                           uint32_t *active_nodes, uint64_t *active_nodes_pointer, uint64_t *num_active)
{
    return guard([&] {
        if (!n || !delta || !out_degree || !active_nodes || !active_nodes_pointer)
            throw Error(ERR_ARG, "null argument or empty graph");
        DevBuf<float> dl(n);
        DevBuf<uint32_t> deg(n), an(n);
        DevBuf<ull> ap(n + 1);
        SUBWAY_CHECK(hipMemcpy(dl.p, delta, n * 4, hipMemcpyHostToDevice));
        SUBWAY_CHECK(hipMemcpy(deg.p, out_degree, n * 4, hipMemcpyHostToDevice));
        ScanScratch s;
        alloc_scan_scratch(s, (n + GEN_TILE - 1) / GEN_TILE);
        struct F { ScanScratch &s; ~F() { free_scan_scratch(s); } } f{s};
        launch_generate_pr(n, dl.p, acc, deg.p, an.p, ap.p, s, nullptr);
        ScanCtl ctl;
        SUBWAY_CHECK(hipMemcpy(&ctl, s.ctl, sizeof(ctl), hipMemcpyDeviceToHost));
        if (ctl.error)
            throw Error(ERR_INTERNAL, std::string("generator: device wait gave up (") + wait_error_name(ctl.error) +
                                          ")");
        SUBWAY_CHECK(hipMemcpy(active_nodes, an.p, ctl.count * 4, hipMemcpyDeviceToHost));
        SUBWAY_CHECK(hipMemcpy(active_nodes_pointer, ap.p, (ctl.count + 1) * 8, hipMemcpyDeviceToHost));
        if (num_active) *num_active = ctl.count;
    });
}

int32_t subway_partition(uint64_t num_active, const uint64_t *active_nodes_pointer, const uint32_t *active_nodes,
                         uint64_t budget_edges, uint64_t max_rows, uint64_t *rows, uint64_t *num_rows)
{
    return guard([&] {
        if (!active_nodes_pointer || !active_nodes || !rows || !num_rows) throw Error(ERR_ARG, "null argument");
        if (!budget_edges) throw Error(ERR_BUDGET, "budget of zero edges");
        *num_rows = 0;
        if (!num_active) return;
        DevBuf<ull> ap(num_active + 1);
        DevBuf<uint32_t> an(num_active);
        SUBWAY_CHECK(hipMemcpy(ap.p, active_nodes_pointer, (num_active + 1) * 8, hipMemcpyHostToDevice));
        SUBWAY_CHECK(hipMemcpy(an.p, active_nodes, num_active * 4, hipMemcpyHostToDevice));
        const uint32_t chunk = 4096;
        DevBuf<PartRow> dr(chunk);
        DevBuf<ull> dout(4);
        std::vector<PartRow> hr(chunk);
        uint64_t start = 0, total = 0;
        for (;;) {
            launch_partition(ap.p, an.p, nullptr, num_active, budget_edges, start, chunk, dr.p, dout.p, nullptr);
            ull out[4];
            SUBWAY_CHECK(hipMemcpy(out, dout.p, sizeof(out), hipMemcpyDeviceToHost));
            SUBWAY_CHECK(hipMemcpy(hr.data(), dr.p, out[0] * sizeof(PartRow), hipMemcpyDeviceToHost));
            for (uint64_t r = 0; r < out[0]; r++, total++) {
                if (total < max_rows) {
                    uint64_t *o = rows + 6 * total;
                    o[0] = hr[r].fromNode;
                    o[1] = hr[r].nodeCount;
                    o[2] = hr[r].fromEdge;
                    o[3] = hr[r].edgeCount;
                    o[4] = hr[r].firstId;
                    o[5] = hr[r].lastId;
                }
            }
            if (out[0] < chunk) break;
            start = out[1];
        }
        *num_rows = total;
    });
}

}  // extern "C"

template <class E>
static void gather_on(const HostCsr<E> &csr, uint64_t A, const uint32_t *nodes, void *out, uint64_t *num_edges)
{
    std::vector<ull> ptr(A + 1);
    uint64_t run = 0;
    for (uint64_t i = 0; i < A; i++) {
        if (nodes[i] >= csr.n) throw Error(ERR_ARG, "active vertex out of range");
        if (i && nodes[i] <= nodes[i - 1]) throw Error(ERR_ARG, "active list must be strictly ascending");
        ptr[i] = run;
        run += csr.degree[nodes[i]];
    }
    ptr[A] = run;
    if (num_edges) *num_edges = run;
    if (!run || !out) return;
    if (!csr.pinned || !csr.d_edges) throw Error(ERR_ARG, "the graph's edge list is not pinned (split-format graph)");
    DevBuf<ull> dp(A + 1), np(csr.n + 1);
    DevBuf<uint32_t> dn(A);
    DevBuf<E> dst(run);
    SUBWAY_CHECK(hipMemcpy(dp.p, ptr.data(), (A + 1) * 8, hipMemcpyHostToDevice));
    SUBWAY_CHECK(hipMemcpy(dn.p, nodes, A * 4, hipMemcpyHostToDevice));
    SUBWAY_CHECK(hipMemcpy(np.p, csr.ptr, (csr.n + 1) * 8, hipMemcpyHostToDevice));
    launch_gather((int)sizeof(E), dn.p, dp.p, 0, A, np.p, csr.d_edges, dst.p, 0, run, nullptr);
    SUBWAY_CHECK(hipDeviceSynchronize());
    SUBWAY_CHECK(hipMemcpy(out, dst.p, run * sizeof(E), hipMemcpyDeviceToHost));
}

extern "C" {

int32_t subway_gather(const subway_graph *g, uint64_t num_active, const uint32_t *active_nodes, void *out,
                      uint64_t *num_edges)
{
    return guard([&] {
        if (!g || (num_active && !active_nodes)) throw Error(ERR_ARG, "null argument");
        if (!num_active) {
            if (num_edges) *num_edges = 0;
            return;
        }
        if (g->weighted) gather_on(*g->w, num_active, active_nodes, out, num_edges);
        else gather_on(*g->u, num_active, active_nodes, out, num_edges);
    });
}

int32_t subway_h2d_peak(uint64_t bytes, int32_t reps, double *gbps)
{
    return guard([&] {
        if (!bytes || reps < 1 || !gbps) throw Error(ERR_ARG, "bad argument");
        void *h = nullptr, *d = nullptr;
        SUBWAY_CHECK(hipHostMalloc(&h, bytes, 0));
        struct Free { void *&h, *&d; ~Free() { if (h) (void)hipHostFree(h); if (d) (void)hipFree(d); } } fr{h, d};
        SUBWAY_CHECK(hipMalloc(&d, bytes));
        std::memset(h, 1, bytes);
        hipStream_t s;
        SUBWAY_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        hipEvent_t a, b;
        SUBWAY_CHECK(hipEventCreate(&a));
        SUBWAY_CHECK(hipEventCreate(&b));
        float best = 1e30f;
        for (int r = 0; r < reps; r++) {
            SUBWAY_CHECK(hipEventRecord(a, s));
            SUBWAY_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
            SUBWAY_CHECK(hipEventRecord(b, s));
            SUBWAY_CHECK(hipEventSynchronize(b));
            float ms = 0;
            SUBWAY_CHECK(hipEventElapsedTime(&ms, a, b));
            best = std::min(best, ms);
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        (void)hipStreamDestroy(s);
        *gbps = (double)bytes / (best * 1e-3) / 1e9;
    });
}

int32_t subway_probe_copy(uint64_t bytes, double *ms)
{
    return guard([&] {
        if (!bytes || bytes % 16) throw Error(ERR_ARG, "size must be a positive multiple of 16");
        DevBuf<uint8_t> a(bytes), b(bytes);
        SUBWAY_CHECK(hipMemset(a.p, 1, bytes));
        hipEvent_t e0, e1;
        SUBWAY_CHECK(hipEventCreate(&e0));
        SUBWAY_CHECK(hipEventCreate(&e1));
        SUBWAY_CHECK(hipDeviceSynchronize());
        SUBWAY_CHECK(hipEventRecord(e0, nullptr));
        launch_copy_probe(a.p, b.p, bytes, nullptr);
        SUBWAY_CHECK(hipEventRecord(e1, nullptr));
        SUBWAY_CHECK(hipEventSynchronize(e1));
        float t = 0;
        SUBWAY_CHECK(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (ms) *ms = t;
    });
}

int32_t subway_mlp_forward(const float *params, const float *x, uint64_t n, float *out)
{
    return guard([&] {
        if (!params || (n && (!x || !out))) throw Error(ERR_ARG, "null argument");
        if (!n) return;
        DevBuf<float> p(MLP_PARAMS), dx(n * MLP_IN), dout(n * MLP_OUT);
        SUBWAY_CHECK(hipMemcpy(p.p, params, MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice));
        SUBWAY_CHECK(hipMemcpy(dx.p, x, n * MLP_IN * sizeof(float), hipMemcpyHostToDevice));
        launch_mlp_forward(p.p, dx.p, n, dout.p, nullptr);
        SUBWAY_CHECK(hipMemcpy(out, dout.p, n * MLP_OUT * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int32_t subway_mlp_grad(const float *params, const float *x, const float *y, const float *mask, uint64_t n,
                        float *grad, float *loss)
{
    return guard([&] {
        if (!params || !grad || !loss || (n && (!x || !y || !mask))) throw Error(ERR_ARG, "null argument");
        DevBuf<float> p(MLP_PARAMS), dx(n * MLP_IN + 1), dy(n * MLP_OUT + 1), dm(n * MLP_OUT + 1);
        DevBuf<float> dg(MLP_PARAMS), dl(1);
        SUBWAY_CHECK(hipMemcpy(p.p, params, MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice));
        if (n) {
            SUBWAY_CHECK(hipMemcpy(dx.p, x, n * MLP_IN * sizeof(float), hipMemcpyHostToDevice));
            SUBWAY_CHECK(hipMemcpy(dy.p, y, n * MLP_OUT * sizeof(float), hipMemcpyHostToDevice));
            SUBWAY_CHECK(hipMemcpy(dm.p, mask, n * MLP_OUT * sizeof(float), hipMemcpyHostToDevice));
        }
        launch_mlp_grad(p.p, dx.p, dy.p, dm.p, n, dg.p, dl.p, nullptr);
        SUBWAY_CHECK(hipMemcpy(grad, dg.p, MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost));
        SUBWAY_CHECK(hipMemcpy(loss, dl.p, sizeof(float), hipMemcpyDeviceToHost));
    });
}

int32_t subway_mlp_train(float *params, const float *x, const float *y, const float *mask, uint64_t n,
                         int32_t steps, float lr, float *loss_per_step)
{
    return guard([&] {
        if (!params || !n || !x || !y || !mask || steps < 0) throw Error(ERR_ARG, "bad argument");
        DevBuf<float> p(MLP_PARAMS), dx(n * MLP_IN), dy(n * MLP_OUT), dm(n * MLP_OUT);
        DevBuf<float> dg(MLP_PARAMS), dl(1), m1(MLP_PARAMS), m2(MLP_PARAMS);
        SUBWAY_CHECK(hipMemcpy(p.p, params, MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice));
        SUBWAY_CHECK(hipMemcpy(dx.p, x, n * MLP_IN * sizeof(float), hipMemcpyHostToDevice));
        SUBWAY_CHECK(hipMemcpy(dy.p, y, n * MLP_OUT * sizeof(float), hipMemcpyHostToDevice));
        SUBWAY_CHECK(hipMemcpy(dm.p, mask, n * MLP_OUT * sizeof(float), hipMemcpyHostToDevice));
        SUBWAY_CHECK(hipMemset(m1.p, 0, MLP_PARAMS * sizeof(float)));
        SUBWAY_CHECK(hipMemset(m2.p, 0, MLP_PARAMS * sizeof(float)));
        std::vector<float> losses(steps > 0 ? steps : 1);
        DevBuf<float> dls(steps > 0 ? steps : 1);
        for (int t = 1; t <= steps; t++) {
            launch_mlp_grad(p.p, dx.p, dy.p, dm.p, n, dg.p, dl.p, nullptr);
            SUBWAY_CHECK(hipMemcpyAsync(dls.p + (t - 1), dl.p, sizeof(float), hipMemcpyDeviceToDevice, nullptr));
            launch_mlp_adam(p.p, dg.p, m1.p, m2.p, t, lr, nullptr);
        }
        SUBWAY_CHECK(hipMemcpy(params, p.p, MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost));
        if (loss_per_step && steps > 0)
            SUBWAY_CHECK(hipMemcpy(loss_per_step, dls.p, steps * sizeof(float), hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
