/*
 * subway_amd.h -- C ABI of the MI355X-native Subway pipeline.
 *
 * The reference has no C ABI: its boundary is a C++ template library linked
 * into ten application binaries (Makefile:12, the subway/ apps).  These entry
 * points expose the same operations with plain pointers and sizes so that a
 * harness (ctypes, cgo, JNI, ...) can drive the path without C++.  Every
 * function returns 0 on success or a negative subway status / positive HIP
 * error code; subway_last_error() returns the message of the calling
 * thread's last failure.  Nothing here exits the process (the reference's
 * gpuErrorcheck does, shared/gpu_error_check.cuh:8-16; the C++ API keeps
 * that behaviour).
 */
#ifndef SUBWAY_AMD_H
#define SUBWAY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (negative); positive values are hipError_t codes */
#define SUBWAY_OK 0
#define SUBWAY_ERR_ARG (-1)
#define SUBWAY_ERR_IO (-2)
#define SUBWAY_ERR_FORMAT (-3)
#define SUBWAY_ERR_BUDGET (-4)
#define SUBWAY_ERR_NOMEM (-5)
#define SUBWAY_ERR_INTERNAL (-6)

/* algorithms and modes (the ten apps are algo x mode) */
#define SUBWAY_BFS 0
#define SUBWAY_CC 1
#define SUBWAY_SSSP 2
#define SUBWAY_SSWP 3
#define SUBWAY_PR 4
#define SUBWAY_SYNC 0
#define SUBWAY_ASYNC 1

/* shared/globals.hpp:27 */
#define SUBWAY_DIST_INFINITY 4294967294u
/* vertex orders of subway_graph_reorder */
#define SUBWAY_ORDER_IN_DEGREE 1

typedef struct subway_graph subway_graph;
typedef struct subway_session subway_session;

typedef struct subway_options {
    int32_t algo;           /* SUBWAY_BFS .. SUBWAY_PR */
    int32_t mode;           /* SUBWAY_SYNC / SUBWAY_ASYNC */
    uint32_t source;        /* --source (default 0) */
    int32_t num_buffers;    /* 1 or 2 device edge buffers (2 overlaps copy/compute) */
    uint64_t budget_bytes;  /* device edge budget; 0 = reference sizing */
    float acc;              /* PR threshold (default 0.01) */
    float init_pr;          /* PR initial delta (default 0.15) */
    int32_t transfer;       /* 0 auto, 1 GPU zero-copy gather, 2 host-thread gather */
    int32_t policy;         /* 0 plain */
    int32_t verbose;        /* print per-partition progress lines */
    int32_t advance_grid;   /* persistent workgroups of the advance kernel; 0 = auto */
    int32_t profile;        /* 1: time every kernel launch and transfer with HIP events */
    int32_t rounds_mode;    /* async: 0 = all rounds of a partition in one persistent launch,
                               1 = one frontier + advance launch pair per round */
    int32_t rounds_grid;    /* workgroups of the persistent rounds kernel; 0 (and the cap) = one below
                               the co-resident maximum per CU */
    int32_t max_rounds;     /* async rounds per partition visit; 0 = 4096 */
    int32_t batch;          /* 1: batched session (BFS/SSSP/SSWP async, <= 64 sources per run) */
    int32_t batch_unroll;   /* batched advance: edges in flight per wave, 1/2/4/8/16/32; 0 = 8 */
    float theta;            /* learned policy without a model: id ranges whose active-edge density is at
                               least theta move whole (default 0.5; > 1 never) */
    uint32_t spin_limit;    /* polls before a device-side wait (scan lookback, grid barrier) gives up and
                               the run fails with SUBWAY_ERR_INTERNAL; 0 = 4194304 */
    int32_t unpacked_edges; /* weighted graphs: 0 (default) moves and buffers each edge as one 32-bit word,
                               destination | weight << ceil(log2 V), when every weight fits the remaining
                               bits; 1 keeps 8-byte {end, w8} edges */
    uint32_t join_window;   /* persistent launches: how long (ticks of the 100 MHz constant clock) the first
                               workgroup waits for the rest of the grid before closing membership; 0 = 50000 */
    int32_t no_resident;    /* 1: move the active edges every global iteration even when the whole graph fits
                               one device buffer (the reference's behaviour); 0 (default): such a graph
                               moves once per run and stays resident, and the learned schedule visits
                               ranges whose whole edge slice a buffer still holds in place */
    uint32_t resident_min_bytes;  /* learned schedule: in-place visits need at least this many bytes of active
                               edges in the resident ranges; 0 = 8 MiB */
    int32_t push_only;      /* BFS on a resident graph: 0 (default) switches between top-down push steps and
                               bottom-up pull steps by frontier size (direction-optimizing BFS); 1 pushes only */
    int32_t no_blocks;      /* 1: a graph with its block layout (subway_graph_build_blocks) still moves its
                               plain edges; 0 (default) moves whole id ranges in the block layout */
    int32_t early_readback; /* single-source runs: 0 (default) on graphs of at least 2^26 vertices, the
                               results are copied to the host once the active edges have fallen to 1/64 of
                               their peak, while the later iterations run, and only the values that change
                               after that are copied at the end; -1: results are copied at the end only;
                               N > 0: the early copy follows the first global iteration whatever the sizes,
                               with a patch list of at most N values */
} subway_options;

typedef struct subway_stats {
    double ms_total;        /* first generate .. values on host */
    double ms_transfer;     /* copy-stream busy time (profile = 1) */
    double ms_advance;      /* advance kernel time (profile = 1) */
    double ms_frontier;     /* frontier kernel time (profile = 1) */
    double ms_generate;     /* generator + partition kernels (profile = 1) */
    double ms_rounds;       /* persistent rounds kernel (profile = 1) */
    uint64_t h2d_bytes;     /* edge bytes moved host -> device */
    uint64_t h2d_dma_bytes;
    uint64_t h2d_gather_bytes;
    uint64_t inner_rounds;  /* frontier launches that selected work */
    uint64_t partitions;
    uint64_t updates;       /* successful relaxations / PR pushes */
    uint64_t advance_launches;
    uint64_t frontier_launches;
    uint64_t transfer_ops;
    uint64_t rounds_launches;
    uint64_t work_edges;    /* edges relaxed over all rounds */
    uint64_t work_entries;  /* vertices that pushed over all rounds */
    uint32_t global_iterations;
    uint32_t trace_len;     /* records written to the trace array */
    uint64_t work_items;    /* partition items scanned by frontier passes over all rounds */
    uint32_t mispredictions;       /* learned policy: iterations whose active edges were > 2x off the prediction */
    uint32_t fallback_iteration;   /* learned policy: first iteration on the plain fallback (0 = none) */
    uint32_t edge_bytes;           /* bytes per edge moved and buffered (4 for packed weighted edges) */
    uint32_t min_members;          /* fewest member workgroups of a persistent launch in the run (0: none) */
    uint64_t lane_work;            /* batched runs: (edge, active source) pairs relaxed over all rounds */
    uint64_t lane_entries;         /* batched runs: active sources of the pushing vertices over all rounds */
    /* the part of work_edges, work_entries and updates done by stand-alone
     * frontier + advance launches (round cap of one, launch-per-round mode,
     * oversized vertices) rather than persistent rounds launches */
    uint64_t pair_work_edges;
    uint64_t pair_work_entries;
    uint64_t pair_updates;
    /* learned schedule: partition visits made in place on a slice a buffer
     * still held, and resident slices skipped because another buffer's
     * slice overlapping them was visited in place in the same iteration */
    uint32_t resident_visits;
    uint32_t resident_overlaps;
    /* direction-optimizing BFS on a resident graph: bottom-up steps taken and
     * the vertices they found */
    uint32_t pull_rounds;
    /* block layout: units moved in it, their bytes and their edges */
    uint32_t blocked_units;
    uint64_t pull_found;
    uint64_t h2d_blocked_bytes;
    uint64_t blocked_edges;
    /* persistent rounds launches over block units: their time (profile = 1),
     * count, and share of work_edges, work_entries and updates */
    double ms_rounds_blocked;
    uint64_t rounds_launches_blocked;
    uint64_t blk_work_edges;
    uint64_t blk_work_entries;
    uint64_t blk_updates;
    /* results: 0 copied to the host at the end of the run; 1 copied after the
     * first global iteration (while the later ones ran), then the values that
     * changed since were patched; 2 copied early, then again in full (more
     * changed than the patch list holds).  values_patched: how many changed. */
    uint64_t early_readback;
    uint64_t values_patched;
} subway_stats;

typedef struct subway_iter_record {
    uint32_t iteration;
    uint32_t partitions;
    uint64_t active_nodes;
    uint64_t active_edges;
    uint64_t h2d_bytes;
    double ms_iteration;
    /* learned policy: the iteration's 12 scheduler features and 3 network heads
     * (log2(next active-edge fraction) / 32, whole-range density threshold,
     * log2(round cap) / 16), the threshold and cap applied, and 1 when the
     * plain schedule ran the iteration */
    float features[12];
    float heads[3];
    float theta;
    uint32_t rounds_cap;
    uint32_t on_plain;
} subway_iter_record;

/* -------- library */
int32_t subway_version(void);
const char *subway_last_error(void);
void subway_default_options(subway_options *opt);
/* number of visible HIP devices (0 without a GPU; never fails) */
int32_t subway_device_count(void);
/* sizes of subway_options, subway_stats and subway_iter_record as this
 * library was built; a binding checks its own layouts against them (no
 * reference counterpart) */
void subway_struct_sizes(uint32_t *options, uint32_t *stats, uint32_t *iter_record);

/* Releases the library's process-wide device resources (the compute/copy
 * stream pair the sessions share) while the HIP runtime is still whole.
 * Call it after every session and graph has been freed, at the end of a
 * program; a session created afterwards makes new streams.  Without it the
 * same happens in an atexit handler.  No reference counterpart (the
 * reference frees nothing, K12). */
void subway_shutdown(void);

/* -------- graphs (host CSR in pinned, device-mapped memory) */
/* Graph<E>(path, isWeighted).ReadGraph(): shared/graph.cu:32-197.
 * .bcsr/.bwcsr (u32 header), .bcsr64/.bwcsr64 (u64 header), .el/.wel text.
 * weighted: 0 no, 1 yes, 2 yes and laid out to move only in the split format
 * (as subway_graph_rmat's 2; .bwcsr files whose weights fit a byte).  A
 * weighted list of 16 GiB or more is read into ordinary memory and then
 * pinned as split arrays when it does not pack into one word per edge and
 * its weights fit a byte, else moved into pinned memory. */
int32_t subway_load(const char *path, int32_t weighted, subway_graph **out);
/* Build from caller arrays: node_pointer[n+1], ends[m], weights[m] or NULL. */
int32_t subway_graph_from_csr(uint64_t n, uint64_t m, const uint64_t *node_pointer, const uint32_t *ends,
                              const uint32_t *weights, subway_graph **out);
/* Reproducible R-MAT (a, b, c; d = 1-a-b-c), self-loops dropped.  weighted:
 * 0 no, 1 yes, 2 yes and laid out as a graph that moves only in the split
 * format (pinned u32 destinations + u8 weights, the 8-byte list in ordinary
 * memory), which large weighted graphs that do not pack get by themselves;
 * needs weights that fit a byte. */
int32_t subway_graph_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                          int32_t weighted, int32_t symmetrize, uint32_t wmin, uint32_t wmax,
                          subway_graph **out);
/* The same R-MAT graph as subway_graph_rmat, byte for byte, built on the
 * GPU (records in generation order, a stable radix sort by source, node
 * pointers from the key boundaries) and copied into pinned host memory: for
 * scales where the host build takes minutes.  Needs about 9 bytes of device
 * memory per generated edge record, twice (SUBWAY_ERR_NOMEM when the device
 * lacks it) and weights that fit a byte.  A graph laid out for the split
 * format (weighted = 2, or large weighted graphs that do not pack) keeps
 * only the split arrays on the host, no 8-byte list.  No reference
 * counterpart (synthetic input). */
int32_t subway_graph_rmat_device(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                                 int32_t weighted, int32_t symmetrize, uint32_t wmin, uint32_t wmax,
                                 subway_graph **out);
/* Zero-copy view of a split-layout graph's host arrays: node_pointer[n+1],
 * destinations ends[m] and byte weights[m] (pinned), valid while the graph
 * lives.  SUBWAY_ERR_ARG when the graph holds no split arrays. */
int32_t subway_graph_split_view(const subway_graph *g, const uint64_t **node_pointer, const uint32_t **ends,
                                const uint8_t **weights);
/* Renumbers the vertices (order SUBWAY_ORDER_IN_DEGREE: descending in-degree,
 * ties by ascending id) on the GPU, in place: the most-targeted vertices'
 * values then share cache lines.  Runs, sessions and batches on the graph
 * still take sources and return values in the original numbering; the
 * accessors (export, split_view, save, subway_generate / partition / gather)
 * see the renumbered graph.  Call it before any session is created on the
 * graph (SUBWAY_ERR_ARG otherwise).  Needs about twice the graph's edge bytes
 * of free device memory.  No counterpart in the reference. */
int32_t subway_graph_reorder(subway_graph *g, int32_t order);
/* Sorts every row by destination and builds the block-compressed copy of
 * the rows (64-edge blocks: an anchor destination, the gaps between
 * consecutive destinations at the block's bit width, weights at the graph's
 * weight width; layout in subway_amd/csrc/blocks.hpp), on the GPU, in
 * place.  Sessions on the graph then move whole id ranges (and compacted
 * units, when smaller) in that layout (fewer bytes over the host link)
 * unless options.no_blocks is set; every result is unchanged (a vertex's
 * out-edges keep their multiset).  The accessors (export, split_view, save)
 * see the sorted rows.  Call it before any session is created on the graph
 * (SUBWAY_ERR_ARG otherwise); needs about 1.5 times the graph's edge bytes
 * plus 13.5 GiB of free device memory.  A process pins a limited amount of
 * host memory: when the words cannot be pinned beside pinned split arrays,
 * those arrays move to ordinary memory first (flag SUBWAY_BLOCKS_UNPIN_SPLIT
 * asks for that up front); compacted units are then gathered by host
 * threads.  No counterpart in the reference (its transfers are the 8-byte
 * edges, shared/subgraph_generator.cu:63-97). */
#define SUBWAY_BLOCKS_UNPIN_SPLIT 1
int32_t subway_graph_build_blocks(subway_graph *g, int32_t flags);
/* The block layout: row_ptr[n+1] (word offset of each row), the words
 * (pinned, num_words of them) and the weight code (bits | base << 6); valid
 * while the graph lives.  SUBWAY_ERR_ARG when the graph has none. */
int32_t subway_graph_blocks_view(const subway_graph *g, const uint64_t **row_ptr, const uint32_t **words,
                                 uint64_t *num_words, uint32_t *weight_code);
/* A graph whose split arrays left pinned memory keeps the split edges of its
 * low-degree id suffix pinned (whole ranges there move in the split format
 * by DMA): *tail_edge = the first edge of that copy, m when there is none.
 * No counterpart in the reference. */
int32_t subway_graph_split_tail(const subway_graph *g, uint64_t *tail_edge);
/* perm[v] = internal id of original vertex v, iperm[p] = original id of
 * internal vertex p (n entries each, valid while the graph lives); both NULL
 * for a graph that keeps its own numbering. */
int32_t subway_graph_order(const subway_graph *g, const uint32_t **perm, const uint32_t **iperm);
/* tools/converter.cpp:88-93,157-162 file layouts; extension picks the variant */
int32_t subway_graph_save(const subway_graph *g, const char *path);
int32_t subway_graph_info(const subway_graph *g, uint64_t *n, uint64_t *m, int32_t *weighted,
                          uint32_t *max_degree, uint32_t *max_degree_vertex);
/* copy the CSR out: node_pointer[n+1], ends[m], weights[m] (may be NULL) */
int32_t subway_graph_export(const subway_graph *g, uint64_t *node_pointer, uint32_t *ends, uint32_t *weights);
void subway_graph_free(subway_graph *g);

/* -------- the path: one application run to its fixpoint.
 * values: n x uint32 (BFS/CC/SSSP/SSWP) or n x float (PR).
 * trace (may be NULL): per-global-iteration records, up to trace_cap. */
int32_t subway_run(const subway_graph *g, const subway_options *opt, void *values, subway_stats *stats,
                   subway_iter_record *trace, uint32_t trace_cap);

/* -------- sessions: the graph's device state and every pipeline buffer
 * built once, then any number of runs (e.g. one query per source).  A
 * session keeps its graph alive; runs on one session are serialised by the
 * caller. */
int32_t subway_session_create(const subway_graph *g, const subway_options *opt, subway_session **out);
int32_t subway_session_run(subway_session *s, uint32_t source, void *values, subway_stats *stats,
                           subway_iter_record *trace, uint32_t trace_cap);
/* The session graph's host value array (graph.value of the reference: u32,
 * or float for PR, one per vertex), holding the last run's results; valid
 * until the next run or the session's end.  Lets a caller read results
 * without the copy `values` makes (pass values = NULL to run). */
const void *subway_session_values(const subway_session *s);
/* Learned policy (options.policy = SUBWAY_POLICY_LEARNED): the scheduler
 * network's SUBWAY_MLP_PARAMS parameters (layout of subway_mlp_forward), or
 * NULL to schedule with the options' fixed theta and round cap.  No
 * counterpart in the reference (its schedule is fixed). */
int32_t subway_session_set_policy(subway_session *s, const float *params);
void subway_session_free(subway_session *s);
/* Batched session (options.batch = 1): k <= 64 sources advance together in
 * one persistent kernel per transferred partition.  rows: n x 64 uint32,
 * vertex-major (rows[v * 64 + i] is the value of v for sources[i]; lanes
 * i >= k hold the initial value).  No reference counterpart: the reference
 * runs k separate processes. */
int32_t subway_session_run_batch(subway_session *s, const uint32_t *sources, int32_t k, uint32_t *rows,
                                 subway_stats *stats, subway_iter_record *trace, uint32_t trace_cap);
/* Autotuner: runs the batch once per candidate wave assignment (4, 8, 16,
 * 32 edges in flight per wave), keeps the fastest; times_ms[4] may be NULL. */
int32_t subway_session_tune_batch(subway_session *s, const uint32_t *sources, int32_t k, int32_t *best_unroll,
                                  double *times_ms);

/* -------- single stages on host arrays (each runs the GPU kernel) */
/* SubgraphGenerator::generate: shared/subgraph_generator.cu:128-162.
 * active_nodes[n], active_nodes_pointer[n+1]. */
int32_t subway_generate(uint64_t n, const uint8_t *label1, const uint8_t *label2, const uint32_t *out_degree,
                        uint32_t *active_nodes, uint64_t *active_nodes_pointer, uint64_t *num_active);
/* No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  subway_generate over arrays the caller already holds in
 * DEVICE memory (label1[n], label2[n], out_degree[n] in; active_nodes[n],
 * active_nodes_pointer[n+1] out); the active count lands in *num_active
 * (host).  Its scratch is a fixed file-scope device array of 4 MiB + 256 B
 * per GPU, held until the process ends, so n is at most 2^32 - 1
 * (SUBWAY_ERR_ARG beyond).  Runs on the null stream and returns after the
 * count is on the host.  Every call on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU race. */
int32_t subway_generate_noalloc(uint64_t n, const uint8_t *label1, const uint8_t *label2, const uint32_t *out_degree,
                                uint32_t *active_nodes, uint64_t *active_nodes_pointer, uint64_t *num_active);
/* SubgraphGenerator::generate(GraphPR&, ..., acc): shared/subgraph_generator.cu:209-243 */
int32_t subway_generate_pr(uint64_t n, const float *delta, float acc, const uint32_t *out_degree,
                           uint32_t *active_nodes, uint64_t *active_nodes_pointer, uint64_t *num_active);
/* Partitioner::partition: shared/partitioner.cu:12-91.  rows: 6 x u64 per
 * partition (fromNode, nodeCount, fromEdge, edgeCount, firstId, lastId). */
int32_t subway_partition(uint64_t num_active, const uint64_t *active_nodes_pointer, const uint32_t *active_nodes,
                         uint64_t budget_edges, uint64_t max_rows, uint64_t *rows, uint64_t *num_rows);
/* Gather of the active vertices' adjacency (the host-thread copy of
 * shared/subgraph_generator.cu:63-97) by the GPU from mapped pinned memory:
 * out receives sum(degree) edges of sizeof(edge) bytes. */
int32_t subway_gather(const subway_graph *g, uint64_t num_active, const uint32_t *active_nodes, void *out,
                      uint64_t *num_edges);

/* -------- the learned scheduler's network (12 -> 64 -> 64 -> 3, FP32)
 * params: SUBWAY_MLP_PARAMS floats = feature mean[12], feature scale[12],
 * W1[12x64], b1[64], W2[64x64], b2[64], W3[64x3], b3[3] (row-major in x out).
 * No reference counterpart. */
#define SUBWAY_MLP_PARAMS 5211
#define SUBWAY_MLP_FEATURES 12
#define SUBWAY_MLP_HEADS 3
/* out[n x 3] = network(x[n x 12]) on the GPU */
int32_t subway_mlp_forward(const float *params, const float *x, uint64_t n, float *out);
/* loss = sum(mask * (out - y)^2) / n and its gradient (grad[SUBWAY_MLP_PARAMS],
 * zero on the normalisation entries), computed on the GPU */
int32_t subway_mlp_grad(const float *params, const float *x, const float *y, const float *mask, uint64_t n,
                        float *grad, float *loss);
/* full-batch Adam for `steps` steps on the GPU; params updated in place;
 * loss_per_step (may be NULL) receives the loss before each step */
int32_t subway_mlp_train(float *params, const float *x, const float *y, const float *mask, uint64_t n,
                         int32_t steps, float lr, float *loss_per_step);

/* -------- measurement */
/* Pinned-copy peak: hipMemcpyAsync of `bytes` from hipHostMalloc memory to
 * the device, best of `reps`, in GB/s (10^9 B/s).  No reference counterpart;
 * the denominator of the achieved host-to-device rate. */
int32_t subway_h2d_peak(uint64_t bytes, int32_t reps, double *gbps);
/* One launch of a plain device-to-device copy kernel that reads and writes
 * `bytes` (multiple of 16): a kernel of known traffic against which hardware
 * byte counters are calibrated.  *ms receives its duration. */
int32_t subway_probe_copy(uint64_t bytes, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* SUBWAY_AMD_H */
