// kernels.hip -- gfx950 kernels of the out-of-memory graph pipeline.
//
//  * select_count / select_scan_chunks / select_emit kernels: the subgraph
//    generator as reduce-then-scan (reference: prePrefix + thrust::reduce +
//    two thrust::exclusive_scan + makeQueue + makeActiveNodesPointer,
//    shared/subgraph_generator.cu:10-60, 128-162).
//  * select_scan_kernel (and scan_tile inside the persistent kernels): one
//    pass of flag -> workgroup scan -> decoupled lookback -> compaction, the
//    per-round frontier filter of every algorithm (the label test/clear at
//    the top of each kernel in shared/gpu_kernels.cu, e.g. :23-28, together
//    with mixLabels/moveUpLabels :515-534).
//  * advance_kernel: edge-balanced relaxation over the frontier's work space
//    (the edge loops of shared/gpu_kernels.cu:42-56 and siblings).
//  * rounds_kernel / rounds_batch_kernel: every async inner round of one
//    resident partition in one persistent launch (shared/gpu_kernels.cu:257-502
//    with the round loop of subway/sssp-async.cu:92-115), single-source and
//    up to 64 sources.
//  * policy_kernel: the learned scheduler's features and fused inference.
//  * partition_kernel: greedy contiguous partitioning of the active list
//    (shared/partitioner.cu:12-91).
//  * gather_kernel: zero-copy gather of active adjacency lists from mapped
//    host memory into the device partition buffer (replaces the host-thread
//    copy of shared/subgraph_generator.cu:63-97 plus the cudaMemcpy of
//    subway/sssp-async.cu:86).
#include "kernels.hpp"
#include "device.hpp"
#include "mlp_device.hpp"
#include "blocks.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <stdexcept>

#include "subway/globals.hpp"
#include "subway/gpu_error_check.hpp"

namespace subway {

using namespace dev;

// Selected-item count of a scan whose totals block is [count, total, error]
// (the layout of ScanCtl); a scan whose lookback gave up selected nothing
// usable, so its consumers see zero items.
__device__ __forceinline__ u64 selected_count(const u64 *totals)
{
    const u64 c = ld_relaxed(&totals[0]);
    return ld_relaxed(&totals[2]) ? 0ull : c;
}

// ====================================================== select-scan tile
// One tile of flag -> workgroup scan -> decoupled lookback -> compaction.
// `lastTile` names the tile that publishes the totals through op.finish.
// Returns false, having emitted nothing, when the lookback gave up.
template <class Op, int BLOCK, int IPT>
__device__ __forceinline__ bool scan_tile(const Op &op, u64 n, u32 tile, u32 epoch, u32 lastTile,
                                          const TileDesc &td, ScanCtl *ctl, const Watch &wt)
{
    constexpr int NW = BLOCK / WAVE;
    __shared__ u32 s_wcnt[NW];
    __shared__ u64 s_wsum[NW];
    __shared__ u32 s_xcnt;
    __shared__ u64 s_xsum;
    __shared__ u32 s_ok;
    const int tid = threadIdx.x;

    const u64 i0 = (u64)tile * (u64)(BLOCK * IPT) + (u64)tid * IPT;
    typename Op::Item it[IPT];
    u64 w[IPT];
    u32 mask = 0;
    op.load(i0, n, it, w, mask);

    u32 cnt = __popc(mask);
    u64 sum = 0;
#pragma unroll
    for (int j = 0; j < IPT; j++)
        if ((mask >> j) & 1u) sum += w[j];

    const int wid = tid / WAVE, lane = lane_id();
    u32 icnt = wave_incl_scan(cnt);
    u64 isum = wave_incl_scan(sum);
    if (lane == WAVE - 1) {
        s_wcnt[wid] = icnt;
        s_wsum[wid] = isum;
    }
    __syncthreads();
    u32 wbc = 0, totc = 0;
    u64 wbs = 0, tots = 0;
#pragma unroll
    for (int q = 0; q < NW; q++) {
        if (q < wid) {
            wbc += s_wcnt[q];
            wbs += s_wsum[q];
        }
        totc += s_wcnt[q];
        tots += s_wsum[q];
    }
    if (wid == 0) {
        u32 xc;
        u64 xs;
        const bool ok = lookback(td, tile, epoch, totc, tots, xc, xs, wt);
        if (lane == 0) {
            s_xcnt = xc;
            s_xsum = xs;
            s_ok = ok ? 1u : 0u;
        }
    }
    __syncthreads();
    if (!s_ok) {
        __syncthreads();
        return false;
    }
    u64 k = (u64)s_xcnt + wbc + (icnt - cnt);
    u64 pre = s_xsum + wbs + (isum - sum);
#pragma unroll
    for (int j = 0; j < IPT; j++) {
        if ((mask >> j) & 1u) {
            op.emit(k, pre, i0 + j, it[j], w[j]);
            k++;
            pre += w[j];
        }
    }
    if (tile == lastTile && tid == 0) {
        op.finish((u64)s_xcnt + totc, s_xsum + tots, ctl);
        atomicAdd((unsigned long long *)&ctl->item_sum, (unsigned long long)n);
    }
    __syncthreads();
    return true;
}

// Standalone launch: one workgroup per tile, tiles claimed in start order.
template <class Op, int BLOCK, int IPT>
__global__ __launch_bounds__(BLOCK) void select_scan_kernel(Op op, u64 n, ScanCtl *ctl,
                                                            TileDesc td, int first, u32 spinLimit)
{
    __shared__ u32 s_tile, s_epoch;
    __shared__ int s_exit;
    const int tid = threadIdx.x;
    if (tid == 0) {
        int ex = (!first && ld_relaxed(&ctl->stop)) ? 1 : 0;
        s_exit = ex;
        if (!ex) {
            s_epoch = ld_relaxed(&ctl->epoch);
            s_tile = fetch_add_relaxed(&ctl->tile_counter, 1u);
        }
    }
    __syncthreads();
    if (s_exit) return;
    const u32 tile = s_tile, epoch = s_epoch;
    (void)scan_tile<Op, BLOCK, IPT>(op, n, tile, epoch, gridDim.x - 1, td, ctl, Watch{&ctl->error, spinLimit});
    // the counters only need the count of finished workgroups; the outputs
    // are consumed by later launches
    if (tid == 0) {
        u32 d = fetch_add_relaxed(&ctl->done_counter, 1u);
        if (d == gridDim.x - 1) {
            st_relaxed(&ctl->done_counter, 0u);
            st_relaxed(&ctl->tile_counter, 0u);
            st_relaxed(&ctl->epoch, epoch + 1);
        }
    }
}

// ================================================ generator: reduce-then-scan
// The generator's selection runs in three wait-free launches: count (each
// workgroup reduces its 16384-vertex chunk to a selected count and degree
// sum, and tallies the degree histogram), scan (one workgroup turns the
// chunk totals into exclusive prefixes and publishes the totals through
// op.finish), emit (each workgroup re-reads its chunk, scans it and
// compacts from its chunk's prefix).  No workgroup waits for another, so the
// time does not grow with a chain of inter-tile waits; the second read of
// the labels and degrees mostly hits L2.

// Exclusive workgroup scan of (count, sum) over SCAN_BLOCK threads; also
// returns the workgroup totals.  Ends with a barrier, so it may be called
// again right away.
__device__ __forceinline__ void block_excl_scan2(u32 cnt, u64 sum, u32 &xc, u64 &xs, u32 &tc, u64 &ts)
{
    constexpr int NW = SCAN_BLOCK / WAVE;
    __shared__ u32 s_c[NW];
    __shared__ u64 s_s[NW];
    const int wid = threadIdx.x / WAVE, lane = lane_id();
    const u32 ic = wave_incl_scan(cnt);
    const u64 is = wave_incl_scan(sum);
    if (lane == WAVE - 1) {
        s_c[wid] = ic;
        s_s[wid] = is;
    }
    __syncthreads();
    u32 bc = 0, totc = 0;
    u64 bs = 0, tots = 0;
#pragma unroll
    for (int q = 0; q < NW; q++) {
        if (q < wid) {
            bc += s_c[q];
            bs += s_s[q];
        }
        totc += s_c[q];
        tots += s_s[q];
    }
    xc = bc + ic - cnt;
    xs = bs + is - sum;
    tc = totc;
    ts = tots;
    __syncthreads();
}

template <class Op>
__global__ __launch_bounds__(SCAN_BLOCK) void select_count_kernel(Op op, u64 n, u64 *aggCnt, u64 *aggSum,
                                                                  ScanCtl *ctl)
{
    static_assert(Op::TALLY, "the count pass re-reads its items; it is for the generator's side-effect-free ops");
    const u64 base = (u64)blockIdx.x * GEN_CHUNK;
    u32 cnt = 0;
    u64 sum = 0;
    u32 h[4] = {0u, 0u, 0u, 0u};
#pragma unroll 1
    for (int r = 0; r < GEN_SUB; r++) {
        const u64 i0 = base + (u64)r * GEN_TILE + (u64)threadIdx.x * GEN_IPT;
        if (i0 >= n) break;
        typename Op::Item it[GEN_IPT];
        u64 w[GEN_IPT];
        u32 mask = 0;
        op.load(i0, n, it, w, mask);
        cnt += __popc(mask);
#pragma unroll
        for (int j = 0; j < GEN_IPT; j++) {
            if ((mask >> j) & 1u) {
                const u64 d = w[j];
                sum += d;
                h[d < 4 ? 0 : d < 32 ? 1 : d < 256 ? 2 : 3]++;
            }
        }
    }
    u32 xc, tc;
    u64 xs, ts;
    block_excl_scan2(cnt, sum, xc, xs, tc, ts);
    __shared__ u32 s_h[4];
    if (threadIdx.x < 4) s_h[threadIdx.x] = 0;
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const u32 t = wave_reduce(h[b]);
        if (lane_id() == 0 && t) atomicAdd(&s_h[b], t);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        aggCnt[blockIdx.x] = tc;
        aggSum[blockIdx.x] = ts;
    }
    if (threadIdx.x < 4 && s_h[threadIdx.x])
        atomicAdd((unsigned long long *)&ctl->hist[threadIdx.x], (unsigned long long)s_h[threadIdx.x]);
}

// One workgroup of 1024 threads: exclusive prefixes of the G chunk totals in
// place, then op.finish with the grand totals.
template <class Op>
__global__ __launch_bounds__(1024) void select_scan_chunks_kernel(Op op, u64 G, u64 *aggCnt, u64 *aggSum,
                                                                  ScanCtl *ctl)
{
    constexpr int T = 1024, NW = T / WAVE;
    __shared__ u64 s_c[NW], s_s[NW];
    const int tid = threadIdx.x, wid = tid / WAVE, lane = lane_id();
    const u64 per = (G + T - 1) / T;
    const u64 b0 = (u64)tid * per < G ? (u64)tid * per : G;
    const u64 b1 = b0 + per < G ? b0 + per : G;
    u64 c = 0, s = 0;
    for (u64 b = b0; b < b1; b++) {
        c += aggCnt[b];
        s += aggSum[b];
    }
    const u64 ic = wave_incl_scan(c), is = wave_incl_scan(s);
    if (lane == WAVE - 1) {
        s_c[wid] = ic;
        s_s[wid] = is;
    }
    __syncthreads();
    u64 bc = 0, bs = 0, tc = 0, ts = 0;
    for (int q = 0; q < NW; q++) {
        if (q < wid) {
            bc += s_c[q];
            bs += s_s[q];
        }
        tc += s_c[q];
        ts += s_s[q];
    }
    u64 xc = bc + ic - c, xs = bs + is - s;
    for (u64 b = b0; b < b1; b++) {
        const u64 cb = aggCnt[b], sb = aggSum[b];
        aggCnt[b] = xc;
        aggSum[b] = xs;
        xc += cb;
        xs += sb;
    }
    if (tid == 0) op.finish(tc, ts, ctl);
}

template <class Op>
__global__ __launch_bounds__(SCAN_BLOCK) void select_emit_kernel(Op op, u64 n, const u64 *aggCnt,
                                                                 const u64 *aggSum)
{
    const u64 base = (u64)blockIdx.x * GEN_CHUNK;
    u64 k0 = aggCnt[blockIdx.x], s0 = aggSum[blockIdx.x];
#pragma unroll 1
    for (int r = 0; r < GEN_SUB; r++) {
        const u64 t0 = base + (u64)r * GEN_TILE;
        if (t0 >= n) break;   // uniform across the workgroup
        const u64 i0 = t0 + (u64)threadIdx.x * GEN_IPT;
        typename Op::Item it[GEN_IPT];
        u64 w[GEN_IPT];
        u32 mask = 0;
        op.load(i0, n, it, w, mask);
        const u32 cnt = __popc(mask);
        u64 sum = 0;
#pragma unroll
        for (int j = 0; j < GEN_IPT; j++)
            if ((mask >> j) & 1u) sum += w[j];
        u32 xc, tc;
        u64 xs, ts;
        block_excl_scan2(cnt, sum, xc, xs, tc, ts);
        u64 k = k0 + xc, pre = s0 + xs;
#pragma unroll
        for (int j = 0; j < GEN_IPT; j++) {
            if ((mask >> j) & 1u) {
                op.emit(k, pre, i0 + j, it[j], w[j]);
                k++;
                pre += w[j];
            }
        }
        k0 += tc;
        s0 += ts;
    }
}

// ------------------------------------------------------- generator ops
struct Empty {};

// activeness = label1 || label2 (shared/subgraph_generator.cu:15)
struct GenLabelOp {
    typedef Empty Item;
    static constexpr bool TALLY = true;
    const uint8_t *l1, *l2;
    const u32 *outDegree;
    u32 *activeNodes;
    u64 *activeNodesPointer;

    __device__ __forceinline__ void load(u64 i0, u64 n, Item *, u64 *w, u32 &mask) const
    {
        if (i0 + GEN_IPT <= n) {
            uint4 a = *reinterpret_cast<const uint4 *>(l1 + i0);
            uint4 b = *reinterpret_cast<const uint4 *>(l2 + i0);
            u32 f[4] = {a.x | b.x, a.y | b.y, a.z | b.z, a.w | b.w};
#pragma unroll
            for (int j = 0; j < GEN_IPT; j++) {
                if ((f[j >> 2] >> ((j & 3) * 8)) & 0xFFu) mask |= 1u << j;
            }
            if (mask) {
#pragma unroll
                for (int q = 0; q < GEN_IPT / 4; q++) {
                    if ((mask >> (4 * q)) & 0xFu) {
                        uint4 d = *reinterpret_cast<const uint4 *>(outDegree + i0 + 4 * q);
                        w[4 * q + 0] = d.x;
                        w[4 * q + 1] = d.y;
                        w[4 * q + 2] = d.z;
                        w[4 * q + 3] = d.w;
                    }
                }
            }
        } else {
            for (int j = 0; j < GEN_IPT; j++) {
                u64 i = i0 + j;
                if (i < n && (l1[i] | l2[i])) {
                    mask |= 1u << j;
                    w[j] = outDegree[i];
                }
            }
        }
    }
    __device__ __forceinline__ void emit(u64 k, u64 pre, u64 i, const Item &, u64) const
    {
        activeNodes[k] = (u32)i;
        activeNodesPointer[k] = pre;
    }
    __device__ __forceinline__ void finish(u64 cnt, u64 sum, ScanCtl *ctl) const
    {
        activeNodesPointer[cnt] = sum;
        st_relaxed(&ctl->count, cnt);
        st_relaxed(&ctl->total, sum);
    }
};

// activeness = delta > acc (shared/subgraph_generator.cu:24-41); the
// residual is FP32 (the stand-alone generator over a caller's array) or FP64
// (the pipeline's own residual)
template <class T>
struct GenPrOp {
    typedef Empty Item;
    static constexpr bool TALLY = true;
    const T *delta;
    T acc;
    const u32 *outDegree;
    u32 *activeNodes;
    u64 *activeNodesPointer;

    __device__ __forceinline__ void load(u64 i0, u64 n, Item *, u64 *w, u32 &mask) const
    {
        if (i0 + GEN_IPT <= n) {
            // 16-byte loads: 4 floats or 2 doubles each
            constexpr int PER = 16 / sizeof(T);
            typedef typename std::conditional<sizeof(T) == 4, float4, double2>::type V;
#pragma unroll
            for (int q = 0; q < GEN_IPT / PER; q++) {
                const V d = *reinterpret_cast<const V *>(delta + i0 + PER * q);
                const T *e = reinterpret_cast<const T *>(&d);
#pragma unroll
                for (int r = 0; r < PER; r++)
                    if (e[r] > acc) mask |= 1u << (PER * q + r);
            }
            if (mask) {
#pragma unroll
                for (int q = 0; q < GEN_IPT / 4; q++) {
                    if ((mask >> (4 * q)) & 0xFu) {
                        uint4 d = *reinterpret_cast<const uint4 *>(outDegree + i0 + 4 * q);
                        w[4 * q + 0] = d.x;
                        w[4 * q + 1] = d.y;
                        w[4 * q + 2] = d.z;
                        w[4 * q + 3] = d.w;
                    }
                }
            }
        } else {
            for (int j = 0; j < GEN_IPT; j++) {
                u64 i = i0 + j;
                if (i < n && delta[i] > acc) {
                    mask |= 1u << j;
                    w[j] = outDegree[i];
                }
            }
        }
    }
    __device__ __forceinline__ void emit(u64 k, u64 pre, u64 i, const Item &, u64) const
    {
        activeNodes[k] = (u32)i;
        activeNodesPointer[k] = pre;
    }
    __device__ __forceinline__ void finish(u64 cnt, u64 sum, ScanCtl *ctl) const
    {
        activeNodesPointer[cnt] = sum;
        st_relaxed(&ctl->count, cnt);
        st_relaxed(&ctl->total, sum);
    }
};

// ------------------------------------------------------- frontier ops
// Statistics of a frontier scan, published by whichever workgroup finishes
// it: inside a persistent launch that is a different workgroup, often on
// another XCD, each round, and the XCDs' L2 caches are not coherent, so the
// counters are device-scope atomics (a plain read-modify-write loses
// increments).  Stand-alone scans (COH false) also count into the pair_*
// words, so the persistent launches' share can be told apart.
template <bool COH>
__device__ __forceinline__ void stat_counts(ScanCtl *ctl, u64 cnt, u64 sum, bool blocked = false)
{
    st_relaxed(&ctl->stop, cnt == 0 ? 1u : 0u);
    if (cnt) atomicAdd((unsigned long long *)&ctl->rounds, 1ull);
    atomicAdd((unsigned long long *)&ctl->entry_sum, (unsigned long long)cnt);
    atomicAdd((unsigned long long *)&ctl->work_sum, (unsigned long long)sum);
    if constexpr (!COH) {
        atomicAdd((unsigned long long *)&ctl->pair_entries, (unsigned long long)cnt);
        atomicAdd((unsigned long long *)&ctl->pair_work, (unsigned long long)sum);
    } else if (blocked) {
        atomicAdd((unsigned long long *)&ctl->blk_entries, (unsigned long long)cnt);
        atomicAdd((unsigned long long *)&ctl->blk_work, (unsigned long long)sum);
    }
}

struct FrItem {
    u32 loc;
    u32 val;
};

template <int CHUNK, bool COH>
__device__ __forceinline__ void write_entry(Entry *entries, u32 *chunkMap, u64 k, u64 pre,
                                            u32 loc, u32 val, u64 deg)
{
    if constexpr (COH) {
        u64 *q = reinterpret_cast<u64 *>(entries + k);
        mst<true>(q, (u64)loc | ((u64)(u32)pre << 32));
        mst<true>(q + 1, (u64)val | ((u64)(u32)deg << 32));
    } else {
        *reinterpret_cast<uint4 *>(entries + k) = make_uint4(loc, (u32)pre, val, (u32)deg);
    }
    // chunk j starts inside [pre, pre + deg): record the entry that holds it
    u64 j = (pre + CHUNK - 1) / CHUNK;
    u64 jend = (pre + deg - 1) / CHUNK;
    for (; j <= jend; j++) mst<COH>(&chunkMap[j], (u32)k);
}

// Entry fields (loc, prefix, val, degree) of entries[k].
template <bool COH>
__device__ __forceinline__ void read_entry(const Entry *entries, u64 k, u32 &loc, u32 &pre, u32 &val, u32 &deg)
{
    if constexpr (COH) {
        const u64 *q = reinterpret_cast<const u64 *>(entries + k);
        const u64 a = mld<true>(q), b = mld<true>(q + 1);
        loc = (u32)a;
        pre = (u32)(a >> 32);
        val = (u32)b;
        deg = (u32)(b >> 32);
    } else {
        const uint4 e = *reinterpret_cast<const uint4 *>(entries + k);
        loc = e.x;
        pre = e.y;
        val = e.z;
        deg = e.w;
    }
}
template <bool COH>
__device__ __forceinline__ void read_entry(const Entry *entries, u64 k, u32 &loc, u32 &pre, u32 &val)
{
    u32 deg;
    read_entry<COH>(entries, k, loc, pre, val, deg);
}

// Labelled vertices of the partition push their current value; their labels
// are cleared as the reference does before its edge loop.  COH: the labels
// and values are exchanged with other workgroups of a persistent launch.
template <bool COH>
struct FrontierLabelOp {
    typedef FrItem Item;
    static constexpr bool TALLY = false;
    const u32 *nodes;
    const u64 *ptr;
    u64 fromEdge;
    uint8_t *labA, *labB;
    const u32 *value;
    Entry *entries;
    u32 *chunkMap;
    u32 idBase;    // items are ids idBase + i when nodes is null (a whole id range)
    const u64 *blk;   // block layout in the buffer: row word offsets (null: plain edges)
    u64 blkBase;      // ... of the buffer's first row
    u32 blkByItem;    // blk is indexed by item (rows gathered in item order), else by vertex id

    __device__ __forceinline__ void load(u64 i0, u64 n, Item *it, u64 *w, u32 &mask) const
    {
#pragma unroll
        for (int j = 0; j < FR_IPT; j++) {
            u64 i = i0 + j;
            if (i < n) {
                u32 id = nodes ? nodes[i] : idBase + (u32)i;
                uint8_t f = mld<COH>(&labA[id]);
                if (labB) f |= mld<COH>(&labB[id]);
                if (f) {
                    mst<COH>(&labA[id], (uint8_t)0);
                    if (labB) mst<COH>(&labB[id], (uint8_t)0);
                    u64 p0 = ptr[i], p1 = ptr[i + 1];
                    if (p1 > p0) {
                        mask |= 1u << j;
                        w[j] = p1 - p0;
                        it[j].loc = blk ? (u32)(blk[blkByItem ? i : (u64)id] - blkBase) : (u32)(p0 - fromEdge);
                        it[j].val = mld<COH>(&value[id]);
                    }
                }
            }
        }
    }
    __device__ __forceinline__ void emit(u64 k, u64 pre, u64, const Item &it, u64 deg) const
    {
        write_entry<ADV_CHUNK, COH>(entries, chunkMap, k, pre, it.loc, it.val, deg);
    }
    __device__ __forceinline__ void finish(u64 cnt, u64 sum, ScanCtl *ctl) const
    {
        st_relaxed(&ctl->count, cnt);
        st_relaxed(&ctl->total, sum);
        stat_counts<COH>(ctl, cnt, sum, blk != nullptr);
    }
};

// PageRank: delta > acc selects; value absorbs delta and the share is
// (float)((d / deg) * 0.85) as in shared/gpu_kernels.cu:233-239.
template <bool COH>
struct FrontierPrOp {
    typedef FrItem Item;
    static constexpr bool TALLY = false;
    const u32 *nodes;
    const u64 *ptr;
    u64 fromEdge;
    double *sum;     // running value sum (FP64)
    double *delta;   // residual (FP64)
    double acc;
    Entry *entries;
    u32 *chunkMap;
    u32 idBase;
    const u64 *blk;   // block layout in the buffer: row word offsets (null: plain edges)
    u64 blkBase;
    u32 blkByItem;    // blk is indexed by item (rows gathered in item order), else by vertex id

    __device__ __forceinline__ void load(u64 i0, u64 n, Item *it, u64 *w, u32 &mask) const
    {
#pragma unroll
        for (int j = 0; j < FR_IPT; j++) {
            u64 i = i0 + j;
            if (i < n) {
                u32 id = nodes ? nodes[i] : idBase + (u32)i;
                const double d = mld<COH>(&delta[id]);
                if (d > acc) {
                    mst<COH>(&sum[id], mld<COH>(&sum[id]) + d);
                    mst<COH>(&delta[id], 0.0);
                    u64 p0 = ptr[i], p1 = ptr[i + 1];
                    if (p1 > p0) {
                        const float share = (float)(d / (double)(p1 - p0) * 0.85);
                        mask |= 1u << j;
                        w[j] = p1 - p0;
                        it[j].loc = blk ? (u32)(blk[blkByItem ? i : (u64)id] - blkBase) : (u32)(p0 - fromEdge);
                        it[j].val = __float_as_uint(share);
                    }
                }
            }
        }
    }
    __device__ __forceinline__ void emit(u64 k, u64 pre, u64, const Item &it, u64 deg) const
    {
        write_entry<ADV_CHUNK, COH>(entries, chunkMap, k, pre, it.loc, it.val, deg);
    }
    __device__ __forceinline__ void finish(u64 cnt, u64 sum, ScanCtl *ctl) const
    {
        st_relaxed(&ctl->count, cnt);
        st_relaxed(&ctl->total, sum);
        stat_counts<COH>(ctl, cnt, sum, blk != nullptr);
    }
};

// ------------------------------------------------- batched (k sources)
// Per-vertex 64-bit activeness masks, one bit per source (lane).
constexpr int BATCH_CHUNK = 256;

// activeness = (mask1 | mask2) != 0
struct GenMaskOp {
    typedef Empty Item;
    static constexpr bool TALLY = true;
    const u64 *m1, *m2;
    const u32 *outDegree;
    u32 *activeNodes;
    u64 *activeNodesPointer;

    __device__ __forceinline__ void load(u64 i0, u64 n, Item *, u64 *w, u32 &mask) const
    {
#pragma unroll
        for (int j = 0; j < GEN_IPT; j++) {
            u64 i = i0 + j;
            if (i < n && (m1[i] | m2[i])) {
                mask |= 1u << j;
                w[j] = outDegree[i];
            }
        }
    }
    __device__ __forceinline__ void emit(u64 k, u64 pre, u64 i, const Item &, u64) const
    {
        activeNodes[k] = (u32)i;
        activeNodesPointer[k] = pre;
    }
    __device__ __forceinline__ void finish(u64 cnt, u64 sum, ScanCtl *ctl) const
    {
        activeNodesPointer[cnt] = sum;
        st_relaxed(&ctl->count, cnt);
        st_relaxed(&ctl->total, sum);
    }
};

struct FrMaskItem {
    u32 loc;
    u32 id;
    u64 m;
};

// A vertex with any source bit set pushes for those sources; its masks are
// cleared (the batched form of the label test/clear).
template <bool COH>
struct FrontierMaskOp {
    typedef FrMaskItem Item;
    static constexpr bool TALLY = false;
    const u32 *nodes;
    const u64 *ptr;
    u64 fromEdge;
    u64 *labA, *labB;
    Entry *entries;
    u64 *entryMask;
    u32 *chunkMap;
    u32 idBase;
    u64 *laneStats;   // ScanCtl lane_work, lane_entries (every thread of the block calls load)

    __device__ __forceinline__ void load(u64 i0, u64 n, Item *it, u64 *w, u32 &mask) const
    {
        u64 lw = 0, lc = 0;
#pragma unroll
        for (int j = 0; j < FR_IPT; j++) {
            u64 i = i0 + j;
            if (i < n) {
                u32 id = nodes ? nodes[i] : idBase + (u32)i;
                u64 m = mld<COH>(&labA[id]);
                if (labB) m |= mld<COH>(&labB[id]);
                if (m) {
                    mst<COH>(&labA[id], 0ull);
                    if (labB) mst<COH>(&labB[id], 0ull);
                    u64 p0 = ptr[i], p1 = ptr[i + 1];
                    if (p1 > p0) {
                        mask |= 1u << j;
                        w[j] = p1 - p0;
                        it[j].loc = (u32)(p0 - fromEdge);
                        it[j].id = id;
                        it[j].m = m;
                        const u64 pc = (u64)__popcll(m);
                        lw += (p1 - p0) * pc;
                        lc += pc;
                    }
                }
            }
        }
        lw = wave_reduce(lw);
        lc = wave_reduce(lc);
        if (lane_id() == 0 && lc) {
            atomicAdd((unsigned long long *)&laneStats[0], (unsigned long long)lw);
            atomicAdd((unsigned long long *)&laneStats[1], (unsigned long long)lc);
        }
    }
    __device__ __forceinline__ void emit(u64 k, u64 pre, u64, const Item &it, u64 deg) const
    {
        write_entry<BATCH_CHUNK, COH>(entries, chunkMap, k, pre, it.loc, it.id, deg);
        mst<COH>(&entryMask[k], it.m);
    }
    __device__ __forceinline__ void finish(u64 cnt, u64 sum, ScanCtl *ctl) const
    {
        st_relaxed(&ctl->count, cnt);
        st_relaxed(&ctl->total, sum);
        stat_counts<COH>(ctl, cnt, sum);
    }
};

// ============================================================ advance
// Update rules of the label algorithms (shared/gpu_kernels.cu): the candidate
// a pushing vertex with value `du` offers across an edge of weight w, the
// improvement test, and the atomic that applies it.
template <int ALGO>
struct Rule;
// cand = value[u] + 1, keep min   (shared/gpu_kernels.cu:42-56)
template <>
struct Rule<ALGO_BFS> {
    __device__ __forceinline__ static u32 cand(u32 du, u32) { return du + 1; }
    __device__ __forceinline__ static bool better(u32 c, u32 old) { return c < old; }
    __device__ __forceinline__ static u32 atomic(u32 *p, u32 c) { return atomicMin(p, c); }
};
// cand = value[u], keep min       (shared/gpu_kernels.cu:90-106)
template <>
struct Rule<ALGO_CC> {
    __device__ __forceinline__ static u32 cand(u32 du, u32) { return du; }
    __device__ __forceinline__ static bool better(u32 c, u32 old) { return c < old; }
    __device__ __forceinline__ static u32 atomic(u32 *p, u32 c) { return atomicMin(p, c); }
};
// cand = value[u] + w, keep min   (shared/gpu_kernels.cu:143-157)
template <>
struct Rule<ALGO_SSSP> {
    __device__ __forceinline__ static u32 cand(u32 du, u32 w) { return du + w; }
    __device__ __forceinline__ static bool better(u32 c, u32 old) { return c < old; }
    __device__ __forceinline__ static u32 atomic(u32 *p, u32 c) { return atomicMin(p, c); }
};
// cand = min(value[u], w), keep max (shared/gpu_kernels.cu:193-207)
template <>
struct Rule<ALGO_SSWP> {
    __device__ __forceinline__ static u32 cand(u32 du, u32 w) { return du < w ? du : w; }
    __device__ __forceinline__ static bool better(u32 c, u32 old) { return c > old; }
    __device__ __forceinline__ static u32 atomic(u32 *p, u32 c) { return atomicMax(p, c); }
};

template <class E>
__device__ __forceinline__ u32 edge_weight(const E &e)
{
    if constexpr (sizeof(E) == 8) return reinterpret_cast<const u32 *>(&e)[1];
    else return 1u;
}

// Storage format -> logical edge of a partition buffer, in two steps: `load`
// issues the memory reads of edge i, `decode` turns what they returned into
// the edge.  The advance issues every load of a phase before decoding any:
// a use of a loaded value makes the compiler wait for that load there, so
// decoding right after each load would serialize the phase.
template <class S>
struct Codec;
template <>
struct Codec<OutEdge> {
    typedef OutEdge T;
    typedef OutEdge Raw;
    __device__ __forceinline__ static Raw load(const OutEdge *p, u64 i, u32) { return p[i]; }
    __device__ __forceinline__ static T decode(const Raw &r, u32) { return r; }
    __device__ __forceinline__ static T get(const OutEdge *p, u64 i, u32 c) { return decode(load(p, i, c), c); }
};
template <>
struct Codec<OutEdgeWeighted> {
    typedef OutEdgeWeighted T;
    typedef OutEdgeWeighted Raw;
    __device__ __forceinline__ static Raw load(const OutEdgeWeighted *p, u64 i, u32) { return p[i]; }
    __device__ __forceinline__ static T decode(const Raw &r, u32) { return r; }
    __device__ __forceinline__ static T get(const OutEdgeWeighted *p, u64 i, u32 c)
    {
        return decode(load(p, i, c), c);
    }
};
template <>
struct Codec<PackedEdge> {
    typedef OutEdgeWeighted T;
    typedef u32 Raw;
    __device__ __forceinline__ static Raw load(const PackedEdge *p, u64 i, u32) { return p[i].bits; }
    // code = destination bits | weight base << 5
    __device__ __forceinline__ static T decode(const Raw &x, u32 code)
    {
        const u32 sh = code & 31u;
        T e;
        e.end = x & ((1u << sh) - 1u);
        e.w8 = (x >> sh) + (code >> 5);
        return e;
    }
    __device__ __forceinline__ static T get(const PackedEdge *p, u64 i, u32 c) { return decode(load(p, i, c), c); }
};

template <>
struct Codec<SplitEdge> {
    typedef OutEdgeWeighted T;
    struct Raw {
        u32 dst;
        u32 w;
    };
    // cap = the buffer's capacity in edges: the weight of edge i is byte i of
    // the array that starts at byte 4 * cap
    __device__ __forceinline__ static Raw load(const SplitEdge *p, u64 i, u32 cap)
    {
        Raw r;
        r.dst = p[i].dst;
        r.w = reinterpret_cast<const uint8_t *>(p)[4ull * cap + i];
        return r;
    }
    __device__ __forceinline__ static T decode(const Raw &r, u32)
    {
        T e;
        e.end = r.dst;
        e.w8 = r.w;
        return e;
    }
    __device__ __forceinline__ static T get(const SplitEdge *p, u64 i, u32 c) { return decode(load(p, i, c), c); }
};

// Relaxes the edges of one work chunk [cbeg, cbeg + ADV_CHUNK) clipped to
// [w0, wend): the chunk's entries are staged in LDS and every position finds
// its entry by binary search.  Once the entries are staged, thread 0 claims
// the caller's next chunk from `claimCtr` (the atomic's latency overlaps the
// edge loads) and stores it to `s_next` (LDS) before the closing barrier;
// ~0u when this chunk's entries were inconsistent.  Returns the successful
// relaxations of the calling thread.
template <int ALGO, class E, bool COH>
__device__ __forceinline__ u32 advance_chunk(u64 j, const Entry *entries, const u32 *chunkMap, const E *edges,
                                             u32 *value, uint8_t *lab, double *delta, u64 count, u64 W, u64 w0,
                                             u64 wend, u64 shift, u32 psh, u32 *err, u32 *claimCtr, u32 *s_next,
                                             u32 nv, u32 wlim)
{
    constexpr bool BLOCKED = std::is_same<E, BlockEdge>::value;
    __shared__ u32 s_pre[ADV_CHUNK + 1];
    __shared__ u32 s_loc[ADV_CHUNK + 1];
    __shared__ u32 s_val[ADV_CHUNK + 1];
    __shared__ u32 s_deg[BLOCKED ? ADV_CHUNK + 1 : 1];
    const int tid = threadIdx.x;
    const u64 cbeg = j * ADV_CHUNK;
    const u64 pbeg = cbeg > w0 ? cbeg : w0;
    const u64 pend = (cbeg + ADV_CHUNK) < wend ? (cbeg + ADV_CHUNK) : wend;
    const u64 k0 = mld<COH>(&chunkMap[j]);
    const u64 k1 = (cbeg + ADV_CHUNK < W) ? mld<COH>(&chunkMap[j + 1]) : count - 1;
    // the chunk's entries are [k0, k1], k1 < count, at most ADV_CHUNK + 1 of
    // them; anything else is frontier output that is not visible (reported,
    // never dereferenced)
    const bool bad = k0 > k1 || k1 >= count || k1 - k0 > (u64)ADV_CHUNK;
    const int nk = bad ? 0 : (int)(k1 - k0 + 1);
    for (int t = tid; t < nk; t += ADV_BLOCK) {
        if constexpr (BLOCKED) read_entry<COH>(entries, k0 + t, s_loc[t], s_pre[t], s_val[t], s_deg[t]);
        else read_entry<COH>(entries, k0 + t, s_loc[t], s_pre[t], s_val[t]);
    }
    if (__syncthreads_or(bad)) {
        if (tid == 0) {
            (void)atomicCAS(err, 0u, WAIT_DATA);
            *s_next = ~0u;
        }
        __syncthreads();
        return 0;
    }
    const u32 next = (tid == 0) ? atomicAdd(claimCtr, 1u) : 0u;
    // every thread owns positions pbeg + tid + u * ADV_BLOCK, u < PPT; the
    // phases below keep PPT edge loads, then PPT value loads, then PPT
    // atomics in flight per thread.  Values a guarded load or atomic returns
    // get no default: they are read only under the same guard, so no merge
    // of the returned value with a default (which makes the compiler wait
    // for the memory operation right there) is needed.
    constexpr int PPT = ADV_CHUNK / ADV_BLOCK;
    static_assert(ADV_CHUNK % ADV_BLOCK == 0, "a chunk is a whole number of positions per thread");
    // block units decode in two halves of the thread's positions (their
    // decode holds more registers per position); plain formats in one step
    constexpr int STEP = BLOCKED ? PPT / 2 : PPT;
    u32 updates = 0;
#pragma unroll
    for (int h = 0; h < PPT; h += STEP) {
        u32 sv[STEP];
        bool ok[STEP];
        u32 dst[STEP], cv[STEP];
        if constexpr (BLOCKED) {
            // block layout (blocks.hpp): the descriptor of the position's
            // block, then the words its gap and weight fields sit in (both
            // words of each field are read; a field that ends in the first
            // one ignores the second, and a buffer keeps two words of slack
            // past its rows)
            const u32 *W = reinterpret_cast<const u32 *>(edges);
            const u32 wbits = psh & 63u, wbase = psh >> BLK_WFIELD;
            constexpr bool WEIGHTS = ALGO == ALGO_SSSP || ALGO == ALGO_SSWP;
            const int lane = lane_id();
            // pay = first payload word of the row, jc = index in the block |
            // edges of the block << 8 | (row of one block) << 16, then the
            // descriptor (da, dc); a row of one block has no second
            // descriptor word, dc is then its first payload word, whose low
            // BLK_WFIELD bits hold the gap width
            u32 pay[STEP], jc[STEP], da[STEP], dc[STEP];
#pragma unroll
            for (int v = 0; v < STEP; v++) {
                const u64 p = pbeg + tid + (u64)(h + v) * ADV_BLOCK;
                ok[v] = p < pend;
                if (ok[v]) {
                    const u32 pp = (u32)p;
                    int lo = 0, hi = nk - 1;
                    while (lo < hi) {
                        int mid = (lo + hi + 1) >> 1;
                        if (s_pre[mid] <= pp) lo = mid;
                        else hi = mid - 1;
                    }
                    const u32 row = s_loc[lo], i = pp - s_pre[lo], d = s_deg[lo];
                    sv[v] = s_val[lo];
                    const u32 k = i >> BLK_SHIFT, left = d - (k << BLK_SHIFT);
                    const u32 nbk = (d + BLK_EDGES - 1u) >> BLK_SHIFT;
                    pay[v] = nbk == 1u ? row + 1u : row + 2u * nbk;
                    jc[v] = (i & (BLK_EDGES - 1u)) | ((left < (u32)BLK_EDGES ? left : (u32)BLK_EDGES) << 8) |
                            (nbk == 1u ? 1u << 16 : 0u);
                    if (row + 2u * k + 1u >= wlim) {
                        // the row's words are not in this buffer: report, never read
                        (void)atomicCAS(err, 0u, WAIT_DATA);
                        ok[v] = false;
                        continue;
                    }
                    da[v] = W[row + 2u * k];
                    dc[v] = W[row + 2u * k + 1u];
                }
            }
            // the gap field's two words (and the weight field's); sh packs
            // the two bit shifts and the gap width.  A wave's 64 lanes hold
            // 64 consecutive positions, filled from lane 0 up; lanes
            // 0 .. j0 - 2 also read gaps 1 .. j0 - 1 of the block lane 0
            // starts in (j0 = lane 0's index in its block): the carry into
            // the wave's first block
            u32 f0[STEP], f1[STEP], g0[STEP], g1[STEP], sh[STEP], c0[STEP], c1[STEP];
#pragma unroll
            for (int v = 0; v < STEP; v++) {
                if (ok[v]) {
                    const u32 jj = jc[v] & 0xFFu, cnt = (jc[v] >> 8) & 0xFFu, one = jc[v] >> 16, wd = dc[v] & 63u;
                    // P = the block's first payload word, base = the bits
                    // in front of its gaps
                    const u32 P = one ? pay[v] : pay[v] + (dc[v] >> BLK_WFIELD);
                    const u32 base = one ? (u32)BLK_WFIELD : 0u;
                    const u32 obit = base + (jj ? (jj - 1u) * wd : 0u);
                    const u32 gbit = base + (cnt - 1u) * wd + jj * wbits;
                    if (P + ((WEIGHTS ? gbit : obit) >> 5) + 1u >= wlim) {
                        (void)atomicCAS(err, 0u, WAIT_DATA);
                        ok[v] = false;
                    } else {
                        f0[v] = W[P + (obit >> 5)];
                        f1[v] = W[P + (obit >> 5) + 1u];
                        u32 gsh = 0;
                        if constexpr (WEIGHTS) {
                            g0[v] = W[P + (gbit >> 5)];
                            g1[v] = W[P + (gbit >> 5) + 1u];
                            gsh = gbit & 31u;
                        }
                        sh[v] = (obit & 31u) | (gsh << 8) | ((jj ? wd : 0u) << 16);
                    }
                }
                if (__builtin_amdgcn_readfirstlane(ok[v] ? 1u : 0u)) {
                    const u32 j0 = __builtin_amdgcn_readfirstlane(jc[v] & 0xFFu);
                    const u32 w0 = __builtin_amdgcn_readfirstlane(dc[v] & 63u);
                    const u32 one0 = __builtin_amdgcn_readfirstlane(jc[v] >> 16);
                    const u32 P0 = __builtin_amdgcn_readfirstlane(one0 ? pay[v] : pay[v] + (dc[v] >> BLK_WFIELD));
                    const u32 base0 = one0 ? (u32)BLK_WFIELD : 0u;
                    if ((u32)lane + 1u < j0) {
                        const u32 cbit = base0 + (u32)lane * w0;
                        if (P0 + (cbit >> 5) + 1u < wlim) {
                            c0[v] = W[P0 + (cbit >> 5)];
                            c1[v] = W[P0 + (cbit >> 5) + 1u];
                        } else {
                            (void)atomicCAS(err, 0u, WAIT_DATA);
                            c0[v] = c1[v] = 0u;
                        }
                    }
                }
            }
#pragma unroll
            for (int v = 0; v < STEP; v++) {
                if (!__builtin_amdgcn_readfirstlane(ok[v] ? 1u : 0u)) continue;   // uniform: no position in the wave
                const u32 j0 = __builtin_amdgcn_readfirstlane(jc[v] & 0xFFu);
                const u32 w0 = __builtin_amdgcn_readfirstlane(dc[v] & 63u);
                const u32 base0 = __builtin_amdgcn_readfirstlane(jc[v] >> 16) ? (u32)BLK_WFIELD : 0u;
                const u32 shadow =
                    ((u32)lane + 1u < j0) ? blk_extract(c0[v], c1[v], (base0 + (u32)lane * w0) & 31u, w0) : 0u;
                const u32 carry = (u32)__builtin_amdgcn_readlane((int)wave_seg_incl_scan(shadow, lane, 0), 63);
                const u32 jj = ok[v] ? (jc[v] & 0xFFu) : 0u;
                const u32 gap = ok[v] && jj ? blk_extract(f0[v], f1[v], sh[v] & 31u, sh[v] >> 16) : 0u;
                // segments: one block's positions in the wave, starting at
                // lane 0 or where the index in the block is 0 (lanes without
                // a position start their own)
                const u64 heads = __ballot(lane == 0 || !ok[v] || jj == 0);
                const u64 upTo = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
                const int segStart = 63 - __clzll((long long)(heads & upTo));
                const u32 S = wave_seg_incl_scan(gap, lane, segStart);
                if (ok[v]) {
                    dst[v] = da[v] + S + (segStart == 0 ? carry : 0u);
                    if (dst[v] >= nv) {
                        // a decoded destination outside the graph: the words
                        // are not this unit's; report, never dereference
                        (void)atomicCAS(err, 0u, WAIT_DATA);
                        ok[v] = false;
                        continue;
                    }
                    if constexpr (ALGO == ALGO_PR) cv[v] = sv[v];
                    else if constexpr (WEIGHTS)
                        cv[v] = Rule<ALGO>::cand(sv[v], wbase + blk_extract(g0[v], g1[v], (sh[v] >> 8) & 31u, wbits));
                    else cv[v] = Rule<ALGO>::cand(sv[v], 1u);
                }
            }
            (void)shift;
        } else {
            typename Codec<E>::Raw raw[STEP];
#pragma unroll
            for (int v = 0; v < STEP; v++) {
                const u64 p = pbeg + tid + (u64)(h + v) * ADV_BLOCK;
                ok[v] = p < pend;
                if (ok[v]) {
                    const u32 pp = (u32)p;
                    int lo = 0, hi = nk - 1;
                    while (lo < hi) {
                        int mid = (lo + hi + 1) >> 1;
                        if (s_pre[mid] <= pp) lo = mid;
                        else hi = mid - 1;
                    }
                    const u64 eidx = (u64)s_loc[lo] + (pp - s_pre[lo]) - shift;
                    raw[v] = Codec<E>::load(edges, eidx, psh);
                    sv[v] = s_val[lo];
                }
            }
#pragma unroll
            for (int v = 0; v < STEP; v++)
                if (ok[v]) {
                    const auto e = Codec<E>::decode(raw[v], psh);
                    dst[v] = e.end;
                    if constexpr (ALGO == ALGO_PR) cv[v] = sv[v];
                    else cv[v] = Rule<ALGO>::cand(sv[v], edge_weight(e));
                }
            (void)nv;
            (void)wlim;
        }
        if constexpr (ALGO == ALGO_PR) {
            // delta[v] += share (shared/gpu_kernels.cu:241-244); the residual is
            // FP64: an FP32 sum at a hub rounds away shares below half its ulp
#pragma unroll
            for (int v = 0; v < STEP; v++)
                if (ok[v]) {
                    unsafeAtomicAdd(&delta[dst[v]], (double)__uint_as_float(cv[v]));
                    updates++;
                }
            (void)value;
            (void)lab;
        } else {
            u32 cur[STEP];
#pragma unroll
            for (int v = 0; v < STEP; v++)
                if (ok[v]) cur[v] = value[dst[v]];
            // every test before the first atomic: a test between two atomics
            // would wait for the earlier atomic as well
            bool tried[STEP];
#pragma unroll
            for (int v = 0; v < STEP; v++) tried[v] = ok[v] && Rule<ALGO>::better(cv[v], cur[v]);
            u32 old[STEP];
#pragma unroll
            for (int v = 0; v < STEP; v++)
                if (tried[v]) old[v] = Rule<ALGO>::atomic(&value[dst[v]], cv[v]);
            // the returned values are taken only after every atomic is issued:
            // the compiler would otherwise fold each improvement test into its
            // atomic's branch and wait for that atomic there
#pragma unroll
            for (int v = 0; v < STEP; v++)
                if (tried[v]) asm volatile("" : "+v"(old[v]));
            // a label is set only where the atomic's old value shows the improvement
#pragma unroll
            for (int v = 0; v < STEP; v++)
                if (tried[v] && Rule<ALGO>::better(cv[v], old[v])) {
                    mst<COH>(&lab[dst[v]], (uint8_t)1);
                    updates++;
                }
            (void)delta;
        }
    }
    if (tid == 0) *s_next = next;
    __syncthreads();
    return updates;
}

template <int ALGO, class E>
__global__ __launch_bounds__(ADV_BLOCK) void advance_kernel(const Entry *entries, const u32 *chunkMap,
                                                            const E *edges, u32 *value, uint8_t *lab,
                                                            double *delta, u64 w0, u64 w1, u64 shift,
                                                            u32 psh, ScanCtl *ctl, u32 nv, u32 wlim)
{
    __shared__ u32 s_chunk;
    const int tid = threadIdx.x;
    // a frontier scan whose wait gave up left no usable entries
    if (ld_relaxed(&ctl->error)) return;
    const u64 count = ctl->count;
    const u64 W = ctl->total;
    const u64 wend = W < w1 ? W : w1;
    if (count == 0 || wend <= w0) return;
    const u64 j0 = w0 / ADV_CHUNK;
    const u64 nchunks = (wend + ADV_CHUNK - 1) / ADV_CHUNK - j0;
    u32 updates = 0;

    // every thread reads the claimed chunk before the next barrier, after
    // which thread 0 may store the following claim
    if (tid == 0) s_chunk = atomicAdd(&ctl->adv_counter, 1u);
    __syncthreads();
    for (u64 c = s_chunk; c < nchunks; c = s_chunk)
        updates += advance_chunk<ALGO, E, false>(j0 + c, entries, chunkMap, edges, value, lab, delta, count, W, w0,
                                                 wend, shift, psh, &ctl->error, &ctl->adv_counter, &s_chunk, nv, wlim);
    u32 wsum = wave_reduce(updates);
    if (lane_id() == 0 && wsum) {
        atomicAdd((unsigned long long *)&ctl->updates, (unsigned long long)wsum);
        atomicAdd((unsigned long long *)&ctl->pair_updates, (unsigned long long)wsum);
    }
    if (tid == 0) {
        u32 d = fetch_add_relaxed(&ctl->adv_done, 1u);
        if (d == gridDim.x - 1) {
            st_relaxed(&ctl->adv_done, 0u);
            st_relaxed(&ctl->adv_counter, 0u);
        }
    }
}

// ================================================== persistent rounds
// ------------------------------------------------ persistent launches
// A persistent launch does not rely on every workgroup of its grid being
// resident at once (kernels of other streams may hold slots).  The first
// workgroup to start waits, at most a window of the constant 100 MHz clock
// (JOIN_WINDOW_TICKS by default), for the rest of the grid, then closes
// membership; workgroups that
// started before the close are the members (ids 0 .. members-1, in start
// order) and take part in every barrier, the others leave at once.  Work is
// claimed through counters, so any member count covers it.
constexpr u32 JOIN_CLOSED = 0x80000000u;
constexpr u64 JOIN_WINDOW_TICKS = 50000;   // 500 us

// Returns false for a workgroup that is not a member (or whose wait for the
// member count gave up; the error word then holds WAIT_BARRIER).
__device__ __forceinline__ bool join_launch(PersistCtl *pc, u32 limit, u32 window, u32 &me, u32 &members)
{
    __shared__ u32 s_me, s_members;
    if (threadIdx.x == 0) {
        u32 old = fetch_add_relaxed(&pc->join, 1u);
        u32 m = 0;
        if (old & JOIN_CLOSED) {
            old = 0;                       // started after the close
        } else if (old == 0) {
            const u64 t0 = wall_clock64();
            while ((ld_relaxed(&pc->join) & ~JOIN_CLOSED) < gridDim.x && wall_clock64() - t0 < window)
                __builtin_amdgcn_s_sleep(1);
            m = __hip_atomic_fetch_or(&pc->join, JOIN_CLOSED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
                ~JOIN_CLOSED;
            st_relaxed(&pc->members, m);
            (void)atomicMin(&pc->min_members, m);
        } else {
            u32 polls = 0;
            while ((m = ld_relaxed(&pc->members)) == 0) {
                if (++polls > limit) {
                    (void)atomicCAS(&pc->error, 0u, WAIT_BARRIER);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        s_me = old;
        s_members = m;
    }
    __syncthreads();
    me = s_me;
    members = s_members;
    return members != 0;
}

// Every workgroup of the grid, member or not, passes here exactly once; the
// last one clears the membership words for the next launch.
__device__ __forceinline__ void leave_launch(PersistCtl *pc)
{
    __syncthreads();
    if (threadIdx.x == 0 && fetch_add_relaxed(&pc->left, 1u) == gridDim.x - 1) {
        st_relaxed(&pc->join, 0u);
        st_relaxed(&pc->members, 0u);
        st_relaxed(&pc->left, 0u);
    }
}

// Two-level barrier over the members.  Everything the workgroups of a
// persistent launch exchange is read and written with agent-scope atomics
// (the COH memory policy), so the barrier orders only their completion: every
// wave waits for its outstanding memory operations, then thread 0 arrives at
// its group counter (up to 8 groups), the last of a group arrives at the top
// counter, and the last overall bumps the generation the others poll.  No L2
// write-back or invalidate is needed (with them a barrier of 768 workgroups
// costs about 60 us instead of 4).  The wait is bounded (`limit` polls,
// WAIT_BARRIER recorded on expiry).  Returns false on every member once any
// wait of the launch has given up (checked after the barrier, so all members
// that pass one barrier agree).
__device__ __forceinline__ bool grid_sync(PersistCtl *pc, u32 &gen, u32 limit, u32 where, u32 me, u32 members)
{
    __shared__ u32 s_fail;
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
        const u32 g = gen;
        const u32 ngroups = members < 8u ? members : 8u;
        const u32 grp = me % ngroups;
        const u32 grpSize = (members - grp + ngroups - 1) / ngroups;
        bool releaser = false;
        u32 a = fetch_add_relaxed(&pc->bar_sub[grp * 16], 1u) + 1;
        if (a == grpSize) {
            // each reset must be performed before the signal that follows
            // it: relaxed operations on different addresses may complete
            // out of order, and a reset landing after the release would
            // erase an arrival at the next barrier (which then never
            // completes)
            st_relaxed(&pc->bar_sub[grp * 16], 0u);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            u32 t = fetch_add_relaxed(&pc->bar_top[0], 1u) + 1;
            if (t == ngroups) {
                st_relaxed(&pc->bar_top[0], 0u);
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                st_relaxed(&pc->bar_gen[0], g + 1);
                releaser = true;
            }
        }
        bool gaveUp = false;
        if (!releaser) {
            u32 polls = 0;
            while (ld_relaxed(&pc->bar_gen[0]) == g) {
                ++polls;
                if (polls > limit || ((polls & 63u) == 0 && ld_relaxed(&pc->error) != 0)) {
                    gaveUp = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        gen = g + 1;
        if (gaveUp && atomicCAS(&pc->error, 0u, WAIT_BARRIER) == 0u) {
            pc->diag[0] = where;
            pc->diag[1] = members;
            pc->diag[2] = ld_relaxed(&pc->bar_sub[grp * 16]);
            pc->diag[3] = ld_relaxed(&pc->bar_top[0]);
            pc->diag[4] = gridDim.x;
        }
        s_fail = (gaveUp || ld_relaxed(&pc->error) != 0) ? 1u : 0u;
    }
    __syncthreads();
    return s_fail == 0;
}

struct RoundsArgs {
    ull n;                 // items of the partition
    const u32 *nodes;      // activeNodes + fromNode
    const ull *ptr;        // activeNodesPointer + fromNode
    ull fromEdge;
    uint8_t *l1, *l2;      // label arrays (label algorithms)
    u32 *value;            // u32 values (label algorithms)
    double *delta;         // PR residual
    double *sum;           // PR running value sum
    double acc;            // PR
    Entry *entries;
    u32 *chunkMap;
    const void *edges;     // device partition buffer
    ScanCtl *sc;
    PersistCtl *pc;
    TileDesc td;
    u32 maxRounds;
    u32 roundBase;         // round index (within the partition) of this launch's first round
    u32 idBase;            // nodes == null: items are the id range idBase ..
    u32 spinLimit;         // poll limit of every wait of the launch
    u32 joinTicks;         // membership window
    u32 packCode;         // EDGE_PACKED: destination bits
    const u64 *blk;        // EDGE_BLOCKS: row word offsets (null otherwise)
    u64 blkBase;           // ... of the buffer's first row
    u32 blkByItem;         // blk indexed by item, else by vertex id
    u32 nv;                // vertices of the graph (decoded destinations are checked against it)
    u32 wlim;              // EDGE_BLOCKS: words of the buffer (no word at or past it is read)
};

// All async inner rounds of one resident partition in one persistent
// launch: phase A selects the labelled vertices (scan tiles claimed by
// counter), grid barrier, phase B relaxes their edges (work chunks claimed by
// counter), grid barrier; the loop ends, uniformly for every workgroup, on
// the first round that selects nothing or after maxRounds rounds.
template <int ALGO, class E>
__device__ __forceinline__ void rounds_body(const RoundsArgs &a, u32 me, u32 members)
{
    __shared__ u32 s_claim, s_gen, s_epoch;
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_gen = ld_relaxed(&a.pc->bar_gen[0]);
        s_epoch = ld_relaxed(&a.sc->epoch);
    }
    __syncthreads();
    u32 gen = s_gen;
    const u32 epoch0 = s_epoch;
    const u32 numTiles = (u32)((a.n + FR_TILE - 1) / FR_TILE);
    const Watch wt{&a.pc->error, a.spinLimit};
    u32 r = 0, conv = 0, updates = 0;
    for (; r < a.maxRounds; r++) {
        const u32 rg = a.roundBase + r;
        const bool even = (rg & 1u) == 0;
        uint8_t *labIn = even ? a.l1 : a.l2;
        uint8_t *labOut = even ? a.l2 : a.l1;
        // phase A: frontier
        for (;;) {
            if (tid == 0) s_claim = atomicAdd(&a.pc->tile_ctr[r & 1], 1u);
            __syncthreads();
            const u32 t = s_claim;
            __syncthreads();
            if (t >= numTiles) break;
            if constexpr (ALGO == ALGO_PR) {
                FrontierPrOp<true> op{a.nodes, a.ptr, a.fromEdge, a.sum, a.delta, a.acc,
                                      a.entries, a.chunkMap, a.idBase, a.blk, a.blkBase, a.blkByItem};
                if (!scan_tile<FrontierPrOp<true>, SCAN_BLOCK, FR_IPT>(op, a.n, t, epoch0 + r, numTiles - 1, a.td, a.sc,
                                                                 wt))
                    break;
            } else {
                FrontierLabelOp<true> op{a.nodes, a.ptr, a.fromEdge, labIn, rg == 0 ? a.l2 : nullptr, a.value,
                                   a.entries, a.chunkMap, a.idBase, a.blk, a.blkBase, a.blkByItem};
                if (!scan_tile<FrontierLabelOp<true>, SCAN_BLOCK, FR_IPT>(op, a.n, t, epoch0 + r, numTiles - 1, a.td,
                                                                    a.sc, wt))
                    break;
            }
        }
        // a wait that gave up leaves the launch's results unusable; every
        // workgroup leaves here and the host reports the error word
        if (!grid_sync(a.pc, gen, a.spinLimit, (r << 1) | 0u, me, members)) return;
        const u64 count = ld_relaxed(&a.sc->count);
        const u64 W = ld_relaxed(&a.sc->total);
        if (count == 0) {
            conv = 1;
            break;
        }
        if (me == 0 && tid == 0) {
            st_relaxed(&a.pc->tile_ctr[(r + 1) & 1], 0u);
            st_relaxed(&a.pc->chunk_ctr[(r + 1) & 1], 0u);
        }
        // phase B: advance
        const u64 nchunks = (W + ADV_CHUNK - 1) / ADV_CHUNK;
        if (tid == 0) s_claim = atomicAdd(&a.pc->chunk_ctr[r & 1], 1u);
        __syncthreads();
        for (u64 c = s_claim; c < nchunks; c = s_claim)
            updates += advance_chunk<ALGO, E, true>(c, a.entries, a.chunkMap, (const E *)a.edges, a.value, labOut,
                                              a.delta, count, W, 0, W, 0, a.packCode, wt.err,
                                              &a.pc->chunk_ctr[r & 1], &s_claim, a.nv, a.wlim);
        if (!grid_sync(a.pc, gen, a.spinLimit, (r << 1) | 1u, me, members)) return;
    }
    u32 wsum = wave_reduce(updates);
    if (lane_id() == 0 && wsum) {
        atomicAdd((unsigned long long *)&a.sc->updates, (unsigned long long)wsum);
        if (a.blk) atomicAdd((unsigned long long *)&a.sc->blk_updates, (unsigned long long)wsum);
    }
    if (me == 0 && tid == 0) {
        st_relaxed(&a.sc->epoch, epoch0 + r + conv);
        a.pc->rounds_done = r;
        a.pc->converged = conv;
        st_relaxed(&a.pc->tile_ctr[0], 0u);
        st_relaxed(&a.pc->tile_ctr[1], 0u);
        st_relaxed(&a.pc->chunk_ctr[0], 0u);
        st_relaxed(&a.pc->chunk_ctr[1], 0u);
    }
}

template <int ALGO, class E>
__global__ __launch_bounds__(ADV_BLOCK) void rounds_kernel(RoundsArgs a)
{
    u32 me, members;
    if (join_launch(a.pc, a.spinLimit, a.joinTicks, me, members)) rounds_body<ALGO, E>(a, me, members);
    leave_launch(a.pc);
}

// ======================================================= batched rounds
// One chunk of BATCH_CHUNK positions: each wave walks a contiguous quarter of
// it, U positions at a time (U edge loads and U destination rows in flight);
// lane s relaxes source s when bit s of the pushing vertex's mask is set.
// dist is vertex-major, 64 u32 per vertex, so a wave's row access is one
// coalesced 256-byte transaction.
template <int ALGO, class E, int U, bool COH>
__device__ __forceinline__ u32 advance_chunk_batch(u64 j, const Entry *entries, const u64 *entryMask,
                                                   const u32 *chunkMap, const E *edges, u32 *dist, u64 *maskOut,
                                                   u64 count, u64 W, u64 w0, u64 wend, u64 shift, u32 psh, u32 *err)
{
    __shared__ u32 b_pre[BATCH_CHUNK + 1];
    __shared__ u32 b_loc[BATCH_CHUNK + 1];
    __shared__ u32 b_id[BATCH_CHUNK + 1];
    __shared__ u64 b_msk[BATCH_CHUNK + 1];
    const int tid = threadIdx.x, lane = lane_id(), wid = tid / WAVE;
    const u64 cfirst = j * BATCH_CHUNK;
    const u64 cbeg = cfirst > w0 ? cfirst : w0;
    const u64 cend = (cfirst + BATCH_CHUNK) < wend ? (cfirst + BATCH_CHUNK) : wend;
    const u64 k0 = mld<COH>(&chunkMap[j]);
    const u64 k1 = (cfirst + BATCH_CHUNK < W) ? mld<COH>(&chunkMap[j + 1]) : count - 1;
    const bool bad = k0 > k1 || k1 >= count || k1 - k0 > (u64)BATCH_CHUNK;
    const int nk = bad ? 0 : (int)(k1 - k0 + 1);
    for (int t = tid; t < nk; t += ADV_BLOCK) {
        read_entry<COH>(entries, k0 + t, b_loc[t], b_pre[t], b_id[t]);
        b_msk[t] = mld<COH>(&entryMask[k0 + t]);
    }
    if (__syncthreads_or(bad)) {
        if (tid == 0) (void)atomicCAS(err, 0u, WAIT_DATA);
        return 0;
    }
    constexpr int NW = ADV_BLOCK / WAVE;
    const u64 span = cend > cbeg ? (cend - cbeg + NW - 1) / NW : 0;
    const u64 q0 = cbeg + (u64)wid * span;
    const u64 q1 = (q0 + span) < cend ? (q0 + span) : cend;
    u32 updates = 0;
    if (q0 < q1) {
        // entry holding q0 (uniform across the wave)
        int lo = 0, hi = nk - 1;
        const u32 pq = (u32)q0;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (b_pre[mid] <= pq) lo = mid;
            else hi = mid - 1;
        }
        // up to 8 positions per step: every load and atomic of a phase is
        // issued before any result is used; at 16 and more the arrays this
        // needs spill under the four-wave register limit, and a spill store
        // of a loaded register waits for the load, so those steps keep the
        // per-position first phase
        if constexpr (U <= 8) {
            int k = lo, kcur = -1;
            u32 du = 0;
            for (u64 p = q0; p < q1; p += U) {
                // phase 1: entries (LDS), raw edge loads and, where the entry
                // changes, the pushing vertex's row; nothing loaded is used
                // yet (guarded loads and atomics get no default value: they
                // are read only under their guard)
                typename Codec<E>::Raw raw[U];
                u32 ldu[U];
                u64 mm[U];
                bool ok[U], fresh[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const u64 pp = p + u;
                    ok[u] = pp < q1;
                    fresh[u] = false;
                    if (ok[u]) {
                        while (k + 1 < nk && b_pre[k + 1] <= (u32)pp) k++;
                        if (k != kcur) {
                            kcur = k;
                            fresh[u] = true;
                            ldu[u] = mld<COH>(&dist[(u64)b_id[k] * WAVE + lane]);
                        }
                        const u64 eidx = (u64)b_loc[k] + ((u32)pp - b_pre[k]) - shift;
                        raw[u] = Codec<E>::load(edges, eidx, psh);
                        mm[u] = b_msk[k];
                    }
                }
                // phase 2: decode, destination rows of the active sources
                u32 vv[U], ww[U], dv[U];
                bool act[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    act[u] = false;
                    if (ok[u]) {
                        const auto e = Codec<E>::decode(raw[u], psh);
                        vv[u] = e.end;
                        ww[u] = edge_weight(e);
                        act[u] = ((mm[u] >> lane) & 1ull) != 0;
                        if (act[u]) dv[u] = dist[(u64)vv[u] * WAVE + lane];
                    }
                }
                // the pushing rows, taken only now that every load of the step
                // is issued (the memory clobber keeps the loads above on this
                // side; otherwise each take moves into its load's branch and
                // waits for that load there)
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (fresh[u]) asm volatile("" : "+v"(ldu[u])::"memory");
                u32 dd[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (fresh[u]) du = ldu[u];
                    dd[u] = du;
                }
                // every test before the first atomic, then all U atomics, then
                // their results (taken after the last atomic is issued)
                u32 cc[U], ol[U];
                bool tr[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    tr[u] = false;
                    if (act[u]) {
                        cc[u] = Rule<ALGO>::cand(dd[u], ww[u]);
                        tr[u] = Rule<ALGO>::better(cc[u], dv[u]);
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (tr[u]) ol[u] = Rule<ALGO>::atomic(&dist[(u64)vv[u] * WAVE + lane], cc[u]);
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (tr[u]) asm volatile("" : "+v"(ol[u]));
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const bool imp = tr[u] && Rule<ALGO>::better(cc[u], ol[u]);
                    const unsigned long long bal = __ballot(imp);
                    if (bal) {
                        if (lane == 0) atomicOr((unsigned long long *)&maskOut[vv[u]], bal);
                        updates += (lane == 0) ? (u32)__popcll(bal) : 0u;
                    }
                }
            }
        } else {
            int k = lo, kcur = -1;
            u32 du = 0;
            for (u64 p = q0; p < q1; p += U) {
                u32 vv[U], ww[U], dd[U], dv[U];
                u64 mm[U];
                bool ok[U];
    #pragma unroll
                for (int u = 0; u < U; u++) {
                    const u64 pp = p + u;
                    ok[u] = pp < q1;
                    mm[u] = 0;
                    if (ok[u]) {
                        while (k + 1 < nk && b_pre[k + 1] <= (u32)pp) k++;
                        if (k != kcur) {
                            kcur = k;
                            du = mld<COH>(&dist[(u64)b_id[k] * WAVE + lane]);
                        }
                        const u64 eidx = (u64)b_loc[k] + ((u32)pp - b_pre[k]) - shift;
                        const auto e = Codec<E>::get(edges, eidx, psh);
                        vv[u] = e.end;
                        ww[u] = edge_weight(e);
                        mm[u] = b_msk[k];
                        dd[u] = du;
                    }
                }
    #pragma unroll
                for (int u = 0; u < U; u++) {
                    dv[u] = 0;
                    if (ok[u] && ((mm[u] >> lane) & 1ull)) dv[u] = dist[(u64)vv[u] * WAVE + lane];
                }
                // all U atomics are issued before any result is used
                u32 cc[U], ol[U];
                bool tr[U];
    #pragma unroll
                for (int u = 0; u < U; u++) {
                    tr[u] = false;
                    cc[u] = 0;
                    if (ok[u] && ((mm[u] >> lane) & 1ull)) {
                        cc[u] = Rule<ALGO>::cand(dd[u], ww[u]);
                        tr[u] = Rule<ALGO>::better(cc[u], dv[u]);
                    }
                    ol[u] = tr[u] ? Rule<ALGO>::atomic(&dist[(u64)vv[u] * WAVE + lane], cc[u]) : 0u;
                }
    #pragma unroll
                for (int u = 0; u < U; u++) {
                    const bool imp = tr[u] && Rule<ALGO>::better(cc[u], ol[u]);
                    const unsigned long long bal = __ballot(imp);
                    if (bal) {
                        if (lane == 0) atomicOr((unsigned long long *)&maskOut[vv[u]], bal);
                        updates += (lane == 0) ? (u32)__popcll(bal) : 0u;
                    }
                }
            }
        }
    }
    __syncthreads();
    return updates;
}

struct BatchArgs {
    ull n;
    const u32 *nodes;
    const ull *ptr;
    ull fromEdge;
    u64 *m1, *m2;          // activeness masks
    u32 *dist;             // [V][64]
    Entry *entries;
    u64 *entryMask;
    u32 *chunkMap;
    const void *edges;
    ScanCtl *sc;
    PersistCtl *pc;
    TileDesc td;
    u32 maxRounds;
    u32 idBase;
    u32 spinLimit;
    u32 joinTicks;         // membership window
    u32 packCode;
};

// All async rounds of one resident partition for up to 64 sources at once,
// in one persistent launch (same phase / barrier structure as rounds_kernel).
template <int ALGO, class E, int U>
__device__ __forceinline__ void rounds_batch_body(const BatchArgs &a, u32 me, u32 members)
{
    __shared__ u32 s_claim, s_gen, s_epoch;
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_gen = ld_relaxed(&a.pc->bar_gen[0]);
        s_epoch = ld_relaxed(&a.sc->epoch);
    }
    __syncthreads();
    u32 gen = s_gen;
    const u32 epoch0 = s_epoch;
    const u32 numTiles = (u32)((a.n + FR_TILE - 1) / FR_TILE);
    const Watch wt{&a.pc->error, a.spinLimit};
    u32 r = 0, conv = 0, updates = 0;
    for (; r < a.maxRounds; r++) {
        const bool even = (r & 1u) == 0;
        u64 *labIn = even ? a.m1 : a.m2;
        u64 *labOut = even ? a.m2 : a.m1;
        for (;;) {
            if (tid == 0) s_claim = atomicAdd(&a.pc->tile_ctr[r & 1], 1u);
            __syncthreads();
            const u32 t = s_claim;
            __syncthreads();
            if (t >= numTiles) break;
            FrontierMaskOp<true> op{a.nodes, a.ptr, a.fromEdge, labIn, r == 0 ? a.m2 : nullptr, a.entries, a.entryMask,
                              a.chunkMap, a.idBase, &a.sc->lane_work};
            if (!scan_tile<FrontierMaskOp<true>, SCAN_BLOCK, FR_IPT>(op, a.n, t, epoch0 + r, numTiles - 1, a.td, a.sc,
                                                               wt))
                break;
        }
        if (!grid_sync(a.pc, gen, a.spinLimit, (r << 1) | 0u, me, members)) return;
        const u64 count = ld_relaxed(&a.sc->count);
        const u64 W = ld_relaxed(&a.sc->total);
        if (count == 0) {
            conv = 1;
            break;
        }
        if (me == 0 && tid == 0) {
            st_relaxed(&a.pc->tile_ctr[(r + 1) & 1], 0u);
            st_relaxed(&a.pc->chunk_ctr[(r + 1) & 1], 0u);
        }
        const u64 nchunks = (W + BATCH_CHUNK - 1) / BATCH_CHUNK;
        for (;;) {
            if (tid == 0) s_claim = atomicAdd(&a.pc->chunk_ctr[r & 1], 1u);
            __syncthreads();
            const u64 c = s_claim;
            __syncthreads();
            if (c >= nchunks) break;
            updates += advance_chunk_batch<ALGO, E, U, true>(c, a.entries, a.entryMask, a.chunkMap, (const E *)a.edges,
                                                       a.dist, labOut, count, W, 0, W, 0, a.packCode, wt.err);
        }
        if (!grid_sync(a.pc, gen, a.spinLimit, (r << 1) | 1u, me, members)) return;
    }
    u32 wsum = wave_reduce(updates);
    if (lane_id() == 0 && wsum) atomicAdd((unsigned long long *)&a.sc->updates, (unsigned long long)wsum);
    if (me == 0 && tid == 0) {
        st_relaxed(&a.sc->epoch, epoch0 + r + conv);
        a.pc->rounds_done = r;
        a.pc->converged = conv;
        st_relaxed(&a.pc->tile_ctr[0], 0u);
        st_relaxed(&a.pc->tile_ctr[1], 0u);
        st_relaxed(&a.pc->chunk_ctr[0], 0u);
        st_relaxed(&a.pc->chunk_ctr[1], 0u);
    }
}

// At most 128 VGPRs (four waves per SIMD): a few registers of the frontier
// scan spill, but the advance, bound by the latency of its row accesses,
// gains more from the extra waves (22% at 64 sources on R-MAT scale 26).
template <int ALGO, class E, int U>
__global__ __launch_bounds__(ADV_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void rounds_batch_kernel(BatchArgs a)
{
    u32 me, members;
    if (join_launch(a.pc, a.spinLimit, a.joinTicks, me, members)) rounds_batch_body<ALGO, E, U>(a, me, members);
    leave_launch(a.pc);
}

// Standalone batched advance over the work window [w0, w1) (a vertex larger
// than a buffer, streamed chunk by chunk).
template <int ALGO, class E>
__global__ __launch_bounds__(ADV_BLOCK) void advance_batch_kernel(const Entry *entries, const u64 *entryMask,
                                                                  const u32 *chunkMap, const E *edges, u32 *dist,
                                                                  u64 *maskOut, u64 w0, u64 w1, u64 shift,
                                                                  u32 psh, ScanCtl *ctl)
{
    // a frontier scan whose wait gave up left no usable entries
    if (ld_relaxed(&ctl->error)) return;
    const u64 count = ctl->count;
    const u64 W = ctl->total;
    const u64 wend = W < w1 ? W : w1;
    if (count == 0 || wend <= w0) return;
    const u64 j0 = w0 / BATCH_CHUNK;
    const u64 j1 = (wend + BATCH_CHUNK - 1) / BATCH_CHUNK;
    u32 updates = 0;
    for (u64 j = j0 + blockIdx.x; j < j1; j += gridDim.x)
        updates += advance_chunk_batch<ALGO, E, 4, false>(j, entries, entryMask, chunkMap, edges, dist, maskOut, count, W,
                                                   w0, wend, shift, psh, &ctl->error);
    u32 wsum = wave_reduce(updates);
    if (lane_id() == 0 && wsum) {
        atomicAdd((unsigned long long *)&ctl->updates, (unsigned long long)wsum);
        atomicAdd((unsigned long long *)&ctl->pair_updates, (unsigned long long)wsum);
    }
}

// Batched initial state: every distance INF (SSWP: 0), masks cleared
// (mask1 all ones when `allActive`, for the first whole-graph generate).
__global__ void init_batch_state_kernel(int algo, u64 n, u32 *dist, u64 *m1, u64 *m2, int allActive)
{
    const u32 init = (algo == ALGO_SSWP) ? 0u : 4294967294u;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n * WAVE; i += (u64)gridDim.x * blockDim.x)
        dist[i] = init;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
        m1[i] = allActive ? ~0ull : 0ull;
        m2[i] = 0ull;
    }
}

// Source s_i gets value 0 (SSWP: INF) in lane i and bit i in mask1.
__global__ void set_batch_sources_kernel(int algo, const u32 *sources, int k, u32 *dist, u64 *m1)
{
    const int i = threadIdx.x;
    if (i < k) {
        const u32 s = sources[i];
        dist[(u64)s * WAVE + i] = (algo == ALGO_SSWP) ? 4294967294u : 0u;
        atomicOr((unsigned long long *)&m1[s], 1ull << i);
    }
}

// ========================================================== partitioner
// Largest k in [lo, hi] with a[k] <= key, given a[lo] <= key and a
// non-decreasing; the whole wave cooperates with a 64-ary search.
__device__ u64 wave_search_le(const u64 *a, u64 lo, u64 hi, u64 key)
{
    const int lane = lane_id();
    while (hi > lo) {
        u64 span = hi - lo;
        u64 step = (span + WAVE - 1) / WAVE;
        u64 q = lo + (u64)(lane + 1) * step;
        if (q > hi) q = hi;
        bool ok = a[q] <= key;
        unsigned long long m = __ballot(ok);
        if (m == 0) {
            u64 q0 = lo + step;
            if (q0 > hi) q0 = hi;
            hi = q0 - 1;
        } else {
            int top = 63 - __clzll((long long)m);
            u64 qtop = __shfl((unsigned long long)q, top, WAVE);
            u64 qnext = (top < WAVE - 1) ? (u64)__shfl((unsigned long long)q, top + 1, WAVE) : hi + 1;
            lo = qtop;
            if (qnext <= hi) hi = qnext - 1;
        }
    }
    return lo;
}

__global__ __launch_bounds__(64) void partition_kernel(const u64 *ptr, const u32 *nodes,
                                                       const u64 *dA, u64 hostA, u64 budget,
                                                       u64 start, u32 maxRows, PartRow *rows,
                                                       u64 *out)
{
    const u64 A = dA ? selected_count(dA) : hostA;
    u64 from = start;
    u32 r = 0;
    while (from < A && r < maxRows) {
        const u64 base = ptr[from];
        const u64 key = base + budget;
        u64 to;
        if (ptr[A] <= key) to = A;
        else to = wave_search_le(ptr, from, A - 1, key);
        if (to == from) to = from + 1;  // vertex larger than the budget
        if (lane_id() == 0) {
            PartRow row;
            row.fromNode = from;
            row.nodeCount = to - from;
            row.fromEdge = base;
            row.edgeCount = ptr[to] - base;
            row.firstId = nodes[from];
            row.lastId = nodes[to - 1];
            rows[r] = row;
        }
        r++;
        from = to;
    }
    if (lane_id() == 0) {
        out[0] = r;
        out[1] = from;
        out[2] = A;
        out[3] = ptr[A];
    }
}

// ========================================================= range stats
// For each static vertex-id range [bounds[i], bounds[i+1]): the slice of
// the active list inside it (lo, hi) and its active edges.
__device__ __forceinline__ u64 lower_bound_u32(const u32 *a, u64 n, u32 key)
{
    u64 lo = 0, hi = n;
    while (lo < hi) {
        u64 mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The scheduler's per-iteration reduction and inference in one workgroup,
// right after the generator: per id range the slice [lo, hi) of the active
// list, its active edges and its first edge offset; the 12 features of the
// iteration; and, when parameters are given, the network's forward pass on
// wave 0.  Clears the generator's degree histogram for the next iteration.
__global__ __launch_bounds__(256) void policy_kernel(PolicyArgs a)
{
    __shared__ u32 s_dense;
    __shared__ float s_x[MLP_IN];
    const int tid = threadIdx.x;
    if (tid == 0) s_dense = 0;
    __syncthreads();
    const u64 A = selected_count(a.totals);
    const u64 EA = A ? ld_relaxed(&a.totals[1]) : 0ull;
    u32 dense = 0;
    for (int i = tid; i < a.P; i += blockDim.x) {
        const u64 lo = lower_bound_u32(a.activeNodes, A, a.bounds[i]);
        const u64 hi = lower_bound_u32(a.activeNodes, A, a.bounds[i + 1]);
        const u64 ea = a.anp[hi] - a.anp[lo];
        a.rstats[4 * i + 0] = lo;
        a.rstats[4 * i + 1] = hi;
        a.rstats[4 * i + 2] = ea;
        a.rstats[4 * i + 3] = a.anp[lo];
        const u64 re = a.nodePointer[a.bounds[i + 1]] - a.nodePointer[a.bounds[i]];
        if (re && 2 * ea >= re) dense++;
    }
    dense = wave_reduce(dense);
    if (lane_id() == 0 && dense) atomicAdd(&s_dense, dense);
    __syncthreads();
    if (tid == 0) {
        const float V = (float)a.V, E = (float)(a.E ? a.E : 1);
        const u64 upd = ld_relaxed(a.updatesNow);
        const u64 lastUpd = upd - a.state->prevUpdates;
        a.state->prevUpdates = upd;
        float hsum = 0.f, h[4];
        for (int b = 0; b < 4; b++) {
            h[b] = (float)a.hist[b];
            hsum += h[b];
        }
        float x[MLP_IN];
        x[0] = log2f(1.f + (float)A) / fmaxf(log2f(V), 1.f);
        x[1] = a.pageRank ? 1.f : 0.f;   // PR pushes every edge of every selected vertex
        x[2] = (float)EA / E;
        x[3] = log2f(1.f + (float)EA) / fmaxf(log2f(E), 1.f);
        for (int b = 0; b < 4; b++) x[4 + b] = hsum > 0.f ? h[b] / hsum : 0.f;
        x[8] = (float)a.bufEdges / E;
        x[9] = (float)a.iteration / 16.f;
        x[10] = log2f(1.f + (float)lastUpd / E) / 8.f;
        x[11] = a.P ? (float)s_dense / (float)a.P : 0.f;
        for (int i = 0; i < MLP_IN; i++) {
            s_x[i] = x[i];
            a.out->feat[i] = x[i];
        }
        for (int b = 0; b < 4; b++) a.hist[b] = 0;
    }
    __syncthreads();
    if (tid < WAVE) {
        float head[MLP_OUT];
        if (a.params) {
            Fwd f;
            forward_wave(a.params, s_x, f);
            for (int m = 0; m < MLP_OUT; m++) head[m] = f.out[m];
        } else {
            for (int m = 0; m < MLP_OUT; m++) head[m] = 0.f;
        }
        if (tid == 0)
            for (int m = 0; m < MLP_OUT; m++) a.out->head[m] = head[m];
    }
}

// ============================================================== gather
// Destination positions [q0, q1) of the partition are mapped to the active
// vertex that owns them through LDS-staged segment starts; sources are read
// straight from mapped pinned host memory.
constexpr int GATHER_SEGS = 1024;
struct EdgeByte {
    uint8_t v;   // one weight of the split format
};

template <class E>
__global__ __launch_bounds__(GATHER_BLOCK) void gather_kernel(const u32 *nodes, const u64 *ptr,
                                                              u64 kbeg, u64 kend, const u64 *nodePointer,
                                                              const E *src, E *dst, u64 edgeBase,
                                                              u64 edgeCount)
{
    __shared__ u64 s_dst[GATHER_SEGS + 1];
    __shared__ u64 s_src[GATHER_SEGS];
    __shared__ u64 s_k[2];
    const int tid = threadIdx.x;
    const u64 nchunks = (edgeCount + GATHER_CHUNK - 1) / GATHER_CHUNK;
    for (u64 chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const u64 q0 = chunk * GATHER_CHUNK;
    const u64 q1 = (q0 + GATHER_CHUNK) < edgeCount ? (q0 + GATHER_CHUNK) : edgeCount;
    if (tid < WAVE) {
        u64 ka = wave_search_le(ptr, kbeg, kend - 1, edgeBase + q0);
        u64 kb = wave_search_le(ptr, ka, kend - 1, edgeBase + q1 - 1);
        if (tid == 0) {
            s_k[0] = ka;
            s_k[1] = kb;
        }
    }
    __syncthreads();
    u64 k0 = s_k[0];
    const u64 k1 = s_k[1];
    while (k0 <= k1) {
        const u64 kk1 = (k1 - k0 + 1 > GATHER_SEGS) ? k0 + GATHER_SEGS - 1 : k1;
        const int ns = (int)(kk1 - k0 + 1);
        for (int t = tid; t < ns; t += GATHER_BLOCK) {
            s_dst[t] = ptr[k0 + t] - edgeBase;
            s_src[t] = nodePointer[nodes[k0 + t]];
        }
        if (tid == 0) s_dst[ns] = ptr[kk1 + 1] - edgeBase;
        __syncthreads();
        u64 pb = s_dst[0] > q0 ? s_dst[0] : q0;
        u64 pe = s_dst[ns] < q1 ? s_dst[ns] : q1;
        constexpr int U = GATHER_CHUNK / GATHER_BLOCK;
        E buf[U];
        u64 at[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            u64 p = pb + (u64)u * GATHER_BLOCK + tid;
            at[u] = ~0ull;
            if (p < pe) {
                int lo = 0, hi = ns - 1;
                while (lo < hi) {
                    int mid = (lo + hi + 1) >> 1;
                    if (s_dst[mid] <= p) lo = mid;
                    else hi = mid - 1;
                }
                at[u] = p;
                buf[u] = src[s_src[lo] + (p - s_dst[lo])];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (at[u] != ~0ull) dst[at[u]] = buf[u];
        __syncthreads();
        k0 = kk1 + 1;
    }
    }
}

// ============================================================== fills
__global__ void fill_u8_kernel(uint8_t *p, uint8_t v, u64 n)
{
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void fill_u32_kernel(u32 *p, u32 v, u64 n)
{
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
        p[i] = v;
}

// Initial vertex state of the label algorithms, written on the device (the
// reference fills host arrays and copies them, e.g. subway/bfs-sync.cu:28-42).
// On a renumbered graph CC starts from the original ids (ids[i], the
// original id of vertex i), so each component still ends at its smallest
// original id.
__global__ void init_label_state_kernel(int algo, int async, u64 n, u32 *value, uint8_t *l1, uint8_t *l2,
                                        u32 source, const u32 *ids)
{
    const u32 INF = 4294967294u;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
        u32 v = (algo == ALGO_CC) ? (ids ? ids[i] : (u32)i) : (algo == ALGO_SSWP ? 0u : INF);
        uint8_t a = async ? 1 : 0;
        uint8_t b = (!async && algo == ALGO_CC) ? 1 : 0;
        if (algo != ALGO_CC && i == source) {
            v = (algo == ALGO_SSWP) ? INF : 0u;
            if (!async) b = 1;
        }
        value[i] = v;
        l1[i] = a;
        l2[i] = b;
    }
}

// PageRank: value 0, delta = initPR (subway/pr-sync.cu:33-40); the running
// sum and the residual are FP64
__global__ void init_pr_state_kernel(u64 n, double *sum, double *delta, double init)
{
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
        sum[i] = 0.0;
        delta[i] = init;
    }
}

// Values of a renumbered graph in the original order: out[v] = in[perm[v]],
// rows of `width` u32 (1 for one value per vertex, 64 for batched rows),
// vertices [v0, v0 + count) into out[0 ..).
__global__ void unpermute_kernel(const u32 *in, const u32 *perm, u64 v0, u64 count, u32 width, u32 *out)
{
    const u64 total = count * width;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (u64)gridDim.x * blockDim.x) {
        const u64 v = i / width, c = i - v * width;
        out[i] = in[(u64)perm[v0 + v] * width + c];
    }
}

// The values that differ from an earlier copy of them: snap[v] holds
// in[v] as it was copied.  Every changed v appends (name[v], in[v]) to
// `list` (name = the original id of a renumbered graph's vertex v; null:
// v itself; the first `cap` entries are written) and counts in *count, one
// atomic per wave.  A lane compares four consecutive values (16-byte loads;
// in and snap 16-byte aligned).
__global__ __launch_bounds__(256) void value_changes_kernel(const u32 *in, const u32 *name, const u32 *snap, u64 n,
                                                            u32 *count, uint2 *list, u32 cap)
{
    const int lane = lane_id();
    const u64 stride = (u64)gridDim.x * blockDim.x;
    const u64 n4 = (n + 3) >> 2;
    // every lane of a wave runs the same number of trips (the wave scan below)
    const u64 trips = (n4 + stride - 1) / stride;
    u64 g = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    for (u64 t = 0; t < trips; t++, g += stride) {
        u32 cur[4] = {0u, 0u, 0u, 0u}, chg = 0;
        const u64 v0 = g << 2;
        if (v0 + 4 <= n) {
            const uint4 a = reinterpret_cast<const uint4 *>(in)[g];
            const uint4 b = reinterpret_cast<const uint4 *>(snap)[g];
            cur[0] = a.x;
            cur[1] = a.y;
            cur[2] = a.z;
            cur[3] = a.w;
            chg = (a.x != b.x ? 1u : 0u) | (a.y != b.y ? 2u : 0u) | (a.z != b.z ? 4u : 0u) | (a.w != b.w ? 8u : 0u);
        } else {
            for (u32 j = 0; j < 4; j++)
                if (v0 + j < n) {
                    cur[j] = in[v0 + j];
                    if (cur[j] != snap[v0 + j]) chg |= 1u << j;
                }
        }
        const u32 nc = (u32)__popc(chg);
        if (!__any(nc != 0)) continue;
        const u32 incl = wave_incl_scan(nc);
        const u32 total = (u32)__shfl((int)incl, WAVE - 1);
        u32 base = 0;
        if (lane == 0) base = atomicAdd(count, total);
        base = (u32)__shfl((int)base, 0) + incl - nc;
        for (u32 j = 0; j < 4; j++)
            if ((chg >> j) & 1u) {
                if (base < cap) list[base] = make_uint2(name ? name[v0 + j] : (u32)(v0 + j), cur[j]);
                base++;
            }
    }
}

// n words from device memory to mapped pinned host memory, written by the
// CUs (16-byte nontemporal stores) rather than by a copy engine: a copy
// engine serves device-to-host copies in order, so one long copy would hold
// back the small read-backs the host waits for between iterations.
typedef u32 u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void store_host_kernel(const u32 *src, u32 *dst, u64 n)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    const u64 t0 = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    const u64 n4 = n >> 2;
    const u32x4 *s4 = reinterpret_cast<const u32x4 *>(src);
    u32x4 *d4 = reinterpret_cast<u32x4 *>(dst);
    for (u64 i = t0; i < n4; i += stride) __builtin_nontemporal_store(s4[i], &d4[i]);
    for (u64 i = (n4 << 2) + t0; i < n; i += stride) dst[i] = src[i];
}

// FP32 result values from the FP64 running sums
__global__ void finish_pr_kernel(u64 n, float *value, const double *sum, float *deltaOut, const double *resid)
{
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
        value[i] = (float)sum[i];
        if (deltaOut) deltaOut[i] = (float)resid[i];
    }
}

// Bottom-up BFS step (direction-optimizing BFS, Beamer, Asanovic and
// Patterson, SC 2012): every unvisited vertex scans its in-edges and stops at
// the first in-neighbour on the frontier (level == cur); found vertices get
// level cur + 1 and the label that makes them the next frontier.  One thread
// per vertex, counts reduced per wave.
__global__ __launch_bounds__(256) void bfs_pull_kernel(u64 n, const u64 *inPtr, const u32 *inAdj, const u32 *outDeg,
                                                       u32 *level, u32 cur, uint8_t *labOut, ScanCtl *ctl)
{
    u64 found = 0, edges = 0;
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (u64)gridDim.x * blockDim.x) {
        if (level[v] != 4294967294u) continue;   // unvisited: DIST_INFINITY
        const u64 e1 = inPtr[v + 1];
        for (u64 e = inPtr[v]; e < e1; e++) {
            if (level[inAdj[e]] == cur) {
                level[v] = cur + 1;
                labOut[v] = 1;
                found++;
                edges += outDeg[v];
                break;
            }
        }
    }
    found = wave_reduce(found);
    edges = wave_reduce(edges);
    if (lane_id() == 0 && found) {
        atomicAdd((unsigned long long *)&ctl->pull_found, (unsigned long long)found);
        atomicAdd((unsigned long long *)&ctl->pull_edges, (unsigned long long)edges);
    }
}

// Plain streaming copy (16 B per lane and iteration); used to calibrate
// hardware byte counters against a kernel of known traffic.
__global__ void copy_probe_kernel(const uint4 *src, uint4 *dst, u64 n16)
{
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (u64)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ============================================================ launchers
static unsigned grid_for(u64 n, int block, unsigned cap = 65535u * 16u)
{
    u64 g = (n + block - 1) / block;
    if (g == 0) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

void alloc_scan_scratch(ScanScratch &s, ull maxTiles)
{
    if (maxTiles == 0) maxTiles = 1;
    SUBWAY_CHECK(hipMalloc(&s.ctl, sizeof(ScanCtl)));
    SUBWAY_CHECK(hipMemset(s.ctl, 0, sizeof(ScanCtl)));
    SUBWAY_CHECK(hipMalloc(&s.td.flag, maxTiles * sizeof(u64)));
    SUBWAY_CHECK(hipMalloc(&s.td.inc, maxTiles * sizeof(u64)));
    // epoch tags start at 0 in the control block; descriptors are cleared to
    // state 3, which no reader accepts, so stale memory never reads as
    // published
    SUBWAY_CHECK(hipMemset(s.td.flag, 0xFF, maxTiles * sizeof(u64)));
    SUBWAY_CHECK(hipMemset(s.td.inc, 0xFF, maxTiles * sizeof(u64)));
    SUBWAY_CHECK(hipDeviceSynchronize());
    s.maxTiles = maxTiles;
    s.epochsIssued = 0;
}

uint32_t scan_wait_error(const ScanScratch &s, hipStream_t st)
{
    uint32_t e = 0;
    SUBWAY_CHECK(hipMemcpyAsync(&e, &s.ctl->error, sizeof(e), hipMemcpyDeviceToHost, st));
    SUBWAY_CHECK(hipStreamSynchronize(st));
    return e;
}

const char *wait_error_name(uint32_t code)
{
    switch (code) {
        case 0: return "none";
        case WAIT_LOOKBACK: return "scan lookback";
        case WAIT_BARRIER: return "grid barrier";
        case WAIT_DATA: return "inconsistent frontier data";
        default: return "unknown wait";
    }
}

void free_scan_scratch(ScanScratch &s)
{
    if (s.ctl) (void)hipFree(s.ctl);
    if (s.td.flag) (void)hipFree(s.td.flag);
    if (s.td.inc) (void)hipFree(s.td.inc);
    s = ScanScratch{};
}

__global__ void fill_u64_kernel(u64 *p, u64 v, u64 n)
{
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) p[i] = v;
}

// Charge `epochs` launch epochs to the scratch; clears the descriptors (with a
// fill kernel, in stream order) before the 22-bit tag could wrap onto a word
// that is still in memory.
static void charge_epochs(ScanScratch &s, u64 epochs, hipStream_t st)
{
    constexpr u64 LIMIT = 1ull << 21;
    if (epochs >= LIMIT) throw Error(ERR_ARG, "too many scan epochs in one launch");
    if (s.epochsIssued + epochs >= LIMIT) {
        const unsigned g = (unsigned)std::min<u64>((s.maxTiles + 255) / 256, 4096);
        hipLaunchKernelGGL(fill_u64_kernel, dim3(g), dim3(256), 0, st, s.td.flag, ~0ull, (u64)s.maxTiles);
        hipLaunchKernelGGL(fill_u64_kernel, dim3(g), dim3(256), 0, st, s.td.inc, ~0ull, (u64)s.maxTiles);
        SUBWAY_CHECK(hipGetLastError());
        s.epochsIssued = 0;
    }
    s.epochsIssued += epochs;
}

static void check_tiles(ScanScratch &s, u64 tiles, hipStream_t st)
{
    charge_epochs(s, 1, st);
    if (tiles > s.maxTiles)
        throw Error(ERR_INTERNAL, "scan launch needs " + std::to_string(tiles) +
                                      " tiles but scratch holds " + std::to_string(s.maxTiles));
    if (tiles > 0x7FFFFFFFull) throw Error(ERR_ARG, "scan launch too large");
}

// Generator selection: count, scan of the chunk totals, emit (see above).
template <class Op>
static void run_generator(const Op &op, ull n, ScanScratch &s, hipStream_t st)
{
    if (n == 0) throw Error(ERR_ARG, "generate on an empty graph");
    const u64 G = (n + GEN_CHUNK - 1) / GEN_CHUNK;
    if (G > s.maxTiles) throw Error(ERR_INTERNAL, "generator scratch too small");
    if (G > 0x7FFFFFFFull) throw Error(ERR_ARG, "generate too large");
    u64 *aggCnt = s.td.flag, *aggSum = s.td.inc;
    hipLaunchKernelGGL((select_count_kernel<Op>), dim3((unsigned)G), dim3(SCAN_BLOCK), 0, st, op, (u64)n, aggCnt,
                       aggSum, s.ctl);
    hipLaunchKernelGGL((select_scan_chunks_kernel<Op>), dim3(1), dim3(1024), 0, st, op, G, aggCnt, aggSum, s.ctl);
    hipLaunchKernelGGL((select_emit_kernel<Op>), dim3((unsigned)G), dim3(SCAN_BLOCK), 0, st, op, (u64)n, aggCnt,
                       aggSum);
    SUBWAY_CHECK(hipGetLastError());
}

// Word offsets of gathered block rows: item k of the active list is vertex
// ids[k], whose row takes rowPtr[ids[k] + 1] - rowPtr[ids[k]] words;
// wpre[k] = the words of items 0 .. k - 1, wpre[n] = all of them.  The row
// lengths are read by one thread per item (the reads are random), then the
// generator's scan runs over them in place: every item is selected, its
// load reads a[i] before its emit writes a[i], and items are disjoint
// between threads.
__global__ void row_words_kernel(const u32 *ids, const u64 *rowPtr, u64 n, u64 *out)
{
    for (u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (u64)gridDim.x * blockDim.x) {
        const u32 v = ids[k];
        out[k] = rowPtr[(u64)v + 1] - rowPtr[v];
    }
}

struct ScanInPlaceOp {
    typedef Empty Item;
    static constexpr bool TALLY = true;
    u64 *a;

    __device__ __forceinline__ void load(u64 i0, u64 n, Item *, u64 *w, u32 &mask) const
    {
        for (int j = 0; j < GEN_IPT; j++) {
            const u64 i = i0 + j;
            if (i < n) {
                mask |= 1u << j;
                w[j] = a[i];
            }
        }
    }
    __device__ __forceinline__ void emit(u64 k, u64 pre, u64, const Item &, u64) const { a[k] = pre; }
    __device__ __forceinline__ void finish(u64 cnt, u64 sum, ScanCtl *) const { a[cnt] = sum; }
};

void launch_row_words(ull n, const uint32_t *ids, const ull *rowPtr, ull *wpre, ScanScratch &s, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(row_words_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, ids, (const u64 *)rowPtr,
                       (u64)n, (u64 *)wpre);
    SUBWAY_CHECK(hipGetLastError());
    ScanInPlaceOp op{(u64 *)wpre};
    run_generator(op, n, s, st);
}

void launch_generate_labels(ull n, const uint8_t *l1, const uint8_t *l2,
                            const uint32_t *outDegree, uint32_t *activeNodes,
                            ull *activeNodesPointer, ScanScratch &s, hipStream_t st)
{
    GenLabelOp op{l1, l2, outDegree, activeNodes, activeNodesPointer};
    run_generator(op, n, s, st);
}

void launch_generate_pr(ull n, const float *delta, float acc, const uint32_t *outDegree,
                        uint32_t *activeNodes, ull *activeNodesPointer, ScanScratch &s,
                        hipStream_t st)
{
    GenPrOp<float> op{delta, acc, outDegree, activeNodes, activeNodesPointer};
    run_generator(op, n, s, st);
}

void launch_generate_pr(ull n, const double *delta, double acc, const uint32_t *outDegree,
                        uint32_t *activeNodes, ull *activeNodesPointer, ScanScratch &s,
                        hipStream_t st)
{
    GenPrOp<double> op{delta, acc, outDegree, activeNodes, activeNodesPointer};
    run_generator(op, n, s, st);
}

void launch_partition(const ull *activeNodesPointer, const uint32_t *activeNodes,
                      const ull *dNumActive, ull numActive, ull budgetEdges,
                      ull start, uint32_t maxRows, PartRow *rows, ull *out,
                      hipStream_t st)
{
    if (budgetEdges == 0) throw Error(ERR_BUDGET, "partition budget of zero edges");
    hipLaunchKernelGGL(partition_kernel, dim3(1), dim3(64), 0, st, activeNodesPointer, activeNodes,
                       (const u64 *)dNumActive, (u64)numActive, (u64)budgetEdges, (u64)start, maxRows,
                       rows, out);
    SUBWAY_CHECK(hipGetLastError());
}

void launch_gather(int edgeBytes, const uint32_t *activeNodes, const ull *activeNodesPointer,
                   ull kbeg, ull kend, const ull *nodePointer, const void *hostEdges,
                   void *dst, ull edgeBase, ull edgeCount, hipStream_t st)
{
    if (edgeCount == 0) return;
    if (kend <= kbeg) throw Error(ERR_ARG, "gather over an empty vertex range");
    // one workgroup per CU loops over the chunks: the slot the compute
    // stream's persistent grids leave free on every CU.  Measured on R-MAT
    // scale 25/26: 128 workgroups gather at half the rate, 512 take slots
    // from persistent launches (474 of 768 members), one per chunk (the
    // former grid) left launches with 2 of 768.
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        SUBWAY_CHECK(hipGetDevice(&dev));
        SUBWAY_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    const u64 chunks = (edgeCount + GATHER_CHUNK - 1) / GATHER_CHUNK;
    const u64 blocks = std::min<u64>(chunks, (u64)std::max(1, cus));
    if (edgeBytes == 8)
        hipLaunchKernelGGL((gather_kernel<OutEdgeWeighted>), dim3((unsigned)blocks), dim3(GATHER_BLOCK), 0,
                           st, activeNodes, activeNodesPointer, (u64)kbeg, (u64)kend, nodePointer,
                           (const OutEdgeWeighted *)hostEdges, (OutEdgeWeighted *)dst, (u64)edgeBase,
                           (u64)edgeCount);
    else if (edgeBytes == 4)
        hipLaunchKernelGGL((gather_kernel<OutEdge>), dim3((unsigned)blocks), dim3(GATHER_BLOCK), 0, st,
                           activeNodes, activeNodesPointer, (u64)kbeg, (u64)kend, nodePointer,
                           (const OutEdge *)hostEdges, (OutEdge *)dst, (u64)edgeBase, (u64)edgeCount);
    else if (edgeBytes == 1)
        hipLaunchKernelGGL((gather_kernel<EdgeByte>), dim3((unsigned)blocks), dim3(GATHER_BLOCK), 0, st,
                           activeNodes, activeNodesPointer, (u64)kbeg, (u64)kend, nodePointer,
                           (const EdgeByte *)hostEdges, (EdgeByte *)dst, (u64)edgeBase, (u64)edgeCount);
    else
        throw Error(ERR_ARG, "gather element size must be 1, 4 or 8 bytes");
    SUBWAY_CHECK(hipGetLastError());
}

void launch_frontier(const FrontierArgs &a, ScanScratch &s, hipStream_t st)
{
    if (a.n == 0) throw Error(ERR_ARG, "frontier over an empty partition");
    u64 tiles = (a.n + FR_TILE - 1) / FR_TILE;
    check_tiles(s, tiles, st);
    if (a.algo == ALGO_PR) {
        FrontierPrOp<false> op{a.nodes, a.ptr, a.fromEdge, a.sum, a.delta, a.acc, a.entries,
                               a.chunkMap, a.idBase, (const u64 *)a.blk, (u64)a.blkBase, a.blkByItem};
        hipLaunchKernelGGL((select_scan_kernel<FrontierPrOp<false>, SCAN_BLOCK, FR_IPT>), dim3((unsigned)tiles),
                           dim3(SCAN_BLOCK), 0, st, op, (u64)a.n, s.ctl, s.td, a.first,
                           s.spinLimit);
    } else {
        FrontierLabelOp<false> op{a.nodes, a.ptr, a.fromEdge, a.labA, a.labB, a.value, a.entries, a.chunkMap,
                                  a.idBase, (const u64 *)a.blk, (u64)a.blkBase, a.blkByItem};
        hipLaunchKernelGGL((select_scan_kernel<FrontierLabelOp<false>, SCAN_BLOCK, FR_IPT>),
                           dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, st, op, (u64)a.n, s.ctl, s.td,
                           a.first, s.spinLimit);
    }
    SUBWAY_CHECK(hipGetLastError());
}

template <int ALGO, class E>
static void adv(const AdvanceArgs &a, ScanCtl *ctl, hipStream_t st)
{
    hipLaunchKernelGGL((advance_kernel<ALGO, E>), dim3((unsigned)a.grid), dim3(ADV_BLOCK), 0, st,
                       a.entries, a.chunkMap, (const E *)a.edges, a.value, a.labelOut, a.delta,
                       (u64)a.w0, (u64)a.w1, (u64)a.shift, (u32)a.packCode, ctl, (u32)a.nv, (u32)a.wlim);
}

// Launch `L<ALGO, storage>` for the buffer's edge format; SSSP and SSWP need
// a weighted format.  The batched kernels take the plain formats only.
#define SUBWAY_BY_FORMAT(L, ALGO, fmt, ...)                                                            \
    do {                                                                                               \
        if ((fmt) == EDGE_BLOCKS) {                                                                    \
            L<ALGO, BlockEdge>(__VA_ARGS__);                                                           \
        } else {                                                                                       \
            SUBWAY_BY_PLAIN_FORMAT(L, ALGO, fmt, __VA_ARGS__);                                         \
        }                                                                                              \
    } while (0)
#define SUBWAY_BY_PLAIN_FORMAT(L, ALGO, fmt, ...)                                                      \
    do {                                                                                               \
        if ((fmt) == EDGE_PACKED) {                                                                    \
            L<ALGO, PackedEdge>(__VA_ARGS__);                                                          \
        } else if ((fmt) == EDGE_SPLIT) {                                                              \
            L<ALGO, SplitEdge>(__VA_ARGS__);                                                           \
        } else if ((fmt) == EDGE_W64) {                                                                \
            L<ALGO, OutEdgeWeighted>(__VA_ARGS__);                                                     \
        } else if constexpr (ALGO != ALGO_SSSP && ALGO != ALGO_SSWP) {                                 \
            if ((fmt) != EDGE_U32) throw Error(ERR_ARG, "unknown edge format");                        \
            L<ALGO, OutEdge>(__VA_ARGS__);                                                             \
        } else {                                                                                       \
            throw Error(ERR_ARG, "this edge format does not carry the weights the algorithm needs");   \
        }                                                                                              \
    } while (0)

void launch_advance(const AdvanceArgs &a, ScanCtl *ctl, hipStream_t st)
{
    if (a.grid <= 0) throw Error(ERR_ARG, "advance grid must be positive");
    switch (a.algo) {
        case ALGO_BFS: SUBWAY_BY_FORMAT(adv, ALGO_BFS, a.edgeFmt, a, ctl, st); break;
        case ALGO_CC: SUBWAY_BY_FORMAT(adv, ALGO_CC, a.edgeFmt, a, ctl, st); break;
        case ALGO_SSSP: SUBWAY_BY_FORMAT(adv, ALGO_SSSP, a.edgeFmt, a, ctl, st); break;
        case ALGO_SSWP: SUBWAY_BY_FORMAT(adv, ALGO_SSWP, a.edgeFmt, a, ctl, st); break;
        case ALGO_PR: SUBWAY_BY_FORMAT(adv, ALGO_PR, a.edgeFmt, a, ctl, st); break;
        default: throw Error(ERR_ARG, "unknown algorithm");
    }
    SUBWAY_CHECK(hipGetLastError());
}

void launch_init_label_state(int algo, int async, ull n, uint32_t *value, uint8_t *l1, uint8_t *l2,
                             uint32_t source, const uint32_t *ids, hipStream_t st)
{
    hipLaunchKernelGGL(init_label_state_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, algo, async,
                       (u64)n, value, l1, l2, source, ids);
    SUBWAY_CHECK(hipGetLastError());
}

void launch_unpermute(const uint32_t *in, const uint32_t *perm, ull v0, ull count, uint32_t width, uint32_t *out,
                      hipStream_t st)
{
    if (count == 0) return;
    hipLaunchKernelGGL(unpermute_kernel, dim3(grid_for(count * width, 256, 16384)), dim3(256), 0, st, in, perm,
                       (u64)v0, (u64)count, width, out);
    SUBWAY_CHECK(hipGetLastError());
}

void launch_store_host(const uint32_t *src, uint32_t *dstMapped, ull n, int grid, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(store_host_kernel, dim3(grid), dim3(256), 0, st, src, dstMapped, (u64)n);
    SUBWAY_CHECK(hipGetLastError());
}

void launch_value_changes(const uint32_t *in, const uint32_t *name, const uint32_t *snap, ull n, uint32_t *count,
                          uint2 *list, uint32_t cap, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(value_changes_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, in, name, snap, (u64)n,
                       count, list, cap);
    SUBWAY_CHECK(hipGetLastError());
}

void launch_init_pr_state(ull n, double *sum, double *delta, double init, hipStream_t st)
{
    hipLaunchKernelGGL(init_pr_state_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, (u64)n, sum,
                       delta, init);
    SUBWAY_CHECK(hipGetLastError());
}

void launch_finish_pr(ull n, float *value, const double *sum, float *deltaOut, const double *resid, hipStream_t st)
{
    hipLaunchKernelGGL(finish_pr_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, (u64)n, value, sum,
                       deltaOut, resid);
    SUBWAY_CHECK(hipGetLastError());
}

void launch_bfs_pull(ull n, const ull *inPtr, const uint32_t *inAdj, const uint32_t *outDeg, uint32_t *level,
                     uint32_t cur, uint8_t *labOut, ScanCtl *ctl, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(bfs_pull_kernel, dim3(grid_for(n, 256, 65536)), dim3(256), 0, st, (u64)n,
                       (const u64 *)inPtr, inAdj, outDeg, level, cur, labOut, ctl);
    SUBWAY_CHECK(hipGetLastError());
}

void launch_copy_probe(const void *src, void *dst, ull bytes, hipStream_t st)
{
    if (bytes % 16) throw Error(ERR_ARG, "probe copy size must be a multiple of 16");
    ull n16 = bytes / 16;
    if (!n16) return;
    hipLaunchKernelGGL(copy_probe_kernel, dim3(grid_for(n16, 256, 8192)), dim3(256), 0, st, (const uint4 *)src,
                       (uint4 *)dst, n16);
    SUBWAY_CHECK(hipGetLastError());
}

// Grid of a persistent kernel: one workgroup slot per CU below the
// co-resident maximum, the slot a kernel of the copy stream (the gather, one
// workgroup per CU) uses, so that normally every workgroup joins.
static int coresident_grid(int perCU, int cus)
{
    return (perCU > 1 ? perCU - 1 : perCU) * cus;
}

template <int ALGO, class E>
static void rounds_fn(const void **f)
{
    *f = reinterpret_cast<const void *>(&rounds_kernel<ALGO, E>);
}

static const void *rounds_fn_for(int algo, int fmt)
{
    const void *f = nullptr;
    switch (algo) {
        case ALGO_BFS: SUBWAY_BY_FORMAT(rounds_fn, ALGO_BFS, fmt, &f); break;
        case ALGO_CC: SUBWAY_BY_FORMAT(rounds_fn, ALGO_CC, fmt, &f); break;
        case ALGO_SSSP: SUBWAY_BY_FORMAT(rounds_fn, ALGO_SSSP, fmt, &f); break;
        case ALGO_SSWP: SUBWAY_BY_FORMAT(rounds_fn, ALGO_SSWP, fmt, &f); break;
        case ALGO_PR: SUBWAY_BY_FORMAT(rounds_fn, ALGO_PR, fmt, &f); break;
        default: throw Error(ERR_ARG, "unknown algorithm");
    }
    return f;
}

int rounds_max_grid(int algo, int edgeFmt)
{
    const void *f = rounds_fn_for(algo, edgeFmt);
    int perCU = 0;
    SUBWAY_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, f, ADV_BLOCK, 0));
    int dev = 0;
    SUBWAY_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SUBWAY_CHECK(hipGetDeviceProperties(&prop, dev));
    return coresident_grid(perCU, prop.multiProcessorCount);
}

void launch_rounds(int algo, int edgeFmt, const RoundsLaunch &l, int grid, hipStream_t st)
{
    if (l.n == 0) throw Error(ERR_ARG, "rounds over an empty partition");
    if (grid <= 0) throw Error(ERR_ARG, "rounds grid must be positive");
    RoundsArgs a;
    a.n = l.n;
    a.nodes = l.nodes;
    a.ptr = l.ptr;
    a.fromEdge = l.fromEdge;
    a.l1 = l.l1;
    a.l2 = l.l2;
    a.value = l.value;
    a.delta = l.delta;
    a.sum = l.sum;
    a.acc = l.acc;
    a.entries = l.entries;
    a.chunkMap = l.chunkMap;
    a.edges = l.edges;
    charge_epochs(*l.scan, (u64)l.maxRounds + 1, st);
    a.sc = l.scan->ctl;
    a.pc = l.pc;
    a.td = l.scan->td;
    a.spinLimit = l.scan->spinLimit;
    a.maxRounds = l.maxRounds;
    a.roundBase = l.roundBase;
    a.idBase = l.idBase;
    a.packCode = l.packCode;
    a.blk = (const u64 *)l.blk;
    a.blkBase = (u64)l.blkBase;
    a.blkByItem = l.blkByItem;
    a.nv = l.nv;
    a.wlim = l.wlim;
    a.joinTicks = l.joinTicks ? l.joinTicks : (u32)JOIN_WINDOW_TICKS;
    void *params[] = {&a};
    // an ordinary launch: the membership protocol (join_launch) makes the
    // barriers correct for whatever part of the grid is resident, and a
    // cooperative launch is refused while another process holds the
    // device's cooperative queue
    SUBWAY_CHECK(hipLaunchKernel(rounds_fn_for(algo, edgeFmt), dim3((unsigned)grid), dim3(ADV_BLOCK), params, 0, st));
}

void launch_generate_masks(ull n, const ull *m1, const ull *m2, const uint32_t *outDegree, uint32_t *activeNodes,
                           ull *activeNodesPointer, ScanScratch &s, hipStream_t st)
{
    GenMaskOp op{m1, m2, outDegree, activeNodes, activeNodesPointer};
    run_generator(op, n, s, st);
}

void launch_init_batch_state(int algo, ull n, uint32_t *dist, ull *m1, ull *m2, int allActive, hipStream_t st)
{
    hipLaunchKernelGGL(init_batch_state_kernel, dim3(grid_for(n * WAVE, 256, 16384)), dim3(256), 0, st, algo,
                       (u64)n, dist, m1, m2, allActive);
    SUBWAY_CHECK(hipGetLastError());
}

void launch_set_batch_sources(int algo, const uint32_t *sources, int k, uint32_t *dist, ull *m1, hipStream_t st)
{
    if (k < 1 || k > 64) throw Error(ERR_ARG, "a batch holds 1..64 sources");
    hipLaunchKernelGGL(set_batch_sources_kernel, dim3(1), dim3(64), 0, st, algo, sources, k, dist, m1);
    SUBWAY_CHECK(hipGetLastError());
}

void launch_frontier_masks(ull n, const uint32_t *nodes, const ull *ptr, ull fromEdge, ull *labA, ull *labB,
                           Entry *entries, ull *entryMask, uint32_t *chunkMap, ScanScratch &s, hipStream_t st,
                           uint32_t idBase)
{
    if (n == 0) throw Error(ERR_ARG, "frontier over an empty partition");
    u64 tiles = (n + FR_TILE - 1) / FR_TILE;
    check_tiles(s, tiles, st);
    FrontierMaskOp<false> op{nodes, ptr, fromEdge, labA, labB, entries, entryMask, chunkMap, idBase, &s.ctl->lane_work};
    hipLaunchKernelGGL((select_scan_kernel<FrontierMaskOp<false>, SCAN_BLOCK, FR_IPT>), dim3((unsigned)tiles),
                       dim3(SCAN_BLOCK), 0, st, op, (u64)n, s.ctl, s.td, 1, s.spinLimit);
    SUBWAY_CHECK(hipGetLastError());
}

struct BatchWindow {
    const Entry *entries;
    const ull *entryMask;
    const uint32_t *chunkMap;
    const void *edges;
    uint32_t *dist;
    ull *maskOut;
    ull w0, w1, shift;
    uint32_t packCode;
    ScanCtl *ctl;
    int grid;
};

template <int ALGO, class E>
static void adv_batch(const BatchWindow &w, hipStream_t st)
{
    hipLaunchKernelGGL((advance_batch_kernel<ALGO, E>), dim3((unsigned)w.grid), dim3(ADV_BLOCK), 0, st, w.entries,
                       w.entryMask, w.chunkMap, (const E *)w.edges, w.dist, w.maskOut, (u64)w.w0, (u64)w.w1,
                       (u64)w.shift, (u32)w.packCode, w.ctl);
}

void launch_advance_batch_window(int algo, int edgeFmt, uint32_t packCode, const Entry *entries,
                                 const ull *entryMask, const uint32_t *chunkMap, const void *edges, uint32_t *dist,
                                 ull *maskOut, ull w0, ull w1, ull shift, ScanCtl *ctl, int grid, hipStream_t st)
{
    if (grid <= 0) throw Error(ERR_ARG, "grid must be positive");
    const BatchWindow w{entries, entryMask, chunkMap, edges, dist, maskOut, w0, w1, shift, packCode, ctl, grid};
    switch (algo) {
        case ALGO_BFS: SUBWAY_BY_PLAIN_FORMAT(adv_batch, ALGO_BFS, edgeFmt, w, st); break;
        case ALGO_SSSP: SUBWAY_BY_PLAIN_FORMAT(adv_batch, ALGO_SSSP, edgeFmt, w, st); break;
        case ALGO_SSWP: SUBWAY_BY_PLAIN_FORMAT(adv_batch, ALGO_SSWP, edgeFmt, w, st); break;
        default: throw Error(ERR_ARG, "batched runs support BFS, SSSP and SSWP");
    }
    SUBWAY_CHECK(hipGetLastError());
}

template <int ALGO, class E, int U>
static const void *batch_fn()
{
    return reinterpret_cast<const void *>(&rounds_batch_kernel<ALGO, E, U>);
}

template <int ALGO, class E>
static void batch_fn_u(const void **f, int unroll)
{
    switch (unroll) {
        case 1: *f = batch_fn<ALGO, E, 1>(); break;
        case 2: *f = batch_fn<ALGO, E, 2>(); break;
        case 4: *f = batch_fn<ALGO, E, 4>(); break;
        case 8: *f = batch_fn<ALGO, E, 8>(); break;
        case 16: *f = batch_fn<ALGO, E, 16>(); break;
        case 32: *f = batch_fn<ALGO, E, 32>(); break;
        default: throw Error(ERR_ARG, "batched unroll must be 1, 2, 4, 8, 16 or 32");
    }
}

static const void *batch_fn_for(int algo, int fmt, int unroll)
{
    const void *f = nullptr;
    switch (algo) {
        case ALGO_BFS: SUBWAY_BY_PLAIN_FORMAT(batch_fn_u, ALGO_BFS, fmt, &f, unroll); break;
        case ALGO_SSSP: SUBWAY_BY_PLAIN_FORMAT(batch_fn_u, ALGO_SSSP, fmt, &f, unroll); break;
        case ALGO_SSWP: SUBWAY_BY_PLAIN_FORMAT(batch_fn_u, ALGO_SSWP, fmt, &f, unroll); break;
        default: throw Error(ERR_ARG, "batched runs support BFS, SSSP and SSWP");
    }
    return f;
}

int batch_max_grid(int algo, int edgeFmt, int unroll)
{
    const void *f = batch_fn_for(algo, edgeFmt, unroll);
    int perCU = 0;
    SUBWAY_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, f, ADV_BLOCK, 0));
    int dev = 0;
    SUBWAY_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SUBWAY_CHECK(hipGetDeviceProperties(&prop, dev));
    return coresident_grid(perCU, prop.multiProcessorCount);
}

void launch_rounds_batch(int algo, int edgeFmt, int unroll, const BatchLaunch &l, int grid, hipStream_t st)
{
    if (l.n == 0) throw Error(ERR_ARG, "rounds over an empty partition");
    if (grid <= 0) throw Error(ERR_ARG, "rounds grid must be positive");
    BatchArgs a;
    a.n = l.n;
    a.nodes = l.nodes;
    a.ptr = l.ptr;
    a.fromEdge = l.fromEdge;
    a.m1 = l.m1;
    a.m2 = l.m2;
    a.dist = l.dist;
    a.entries = l.entries;
    a.entryMask = l.entryMask;
    a.chunkMap = l.chunkMap;
    a.edges = l.edges;
    charge_epochs(*l.scan, (u64)l.maxRounds + 1, st);
    a.sc = l.scan->ctl;
    a.pc = l.pc;
    a.td = l.scan->td;
    a.spinLimit = l.scan->spinLimit;
    a.maxRounds = l.maxRounds;
    a.idBase = l.idBase;
    a.packCode = l.packCode;
    a.joinTicks = l.joinTicks ? l.joinTicks : (u32)JOIN_WINDOW_TICKS;
    void *params[] = {&a};
    SUBWAY_CHECK(hipLaunchKernel(batch_fn_for(algo, edgeFmt, unroll), dim3((unsigned)grid), dim3(ADV_BLOCK), params, 0,
                                 st));
}

void launch_policy(const PolicyArgs &a, hipStream_t st)
{
    if (a.P <= 0) throw Error(ERR_ARG, "policy over no id ranges");
    hipLaunchKernelGGL(policy_kernel, dim3(1), dim3(256), 0, st, a);
    SUBWAY_CHECK(hipGetLastError());
}

void launch_fill_u8(uint8_t *p, uint8_t v, ull n, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(fill_u8_kernel, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, p, v, (u64)n);
    SUBWAY_CHECK(hipGetLastError());
}
void launch_fill_u32(uint32_t *p, uint32_t v, ull n, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, p, v, (u64)n);
    SUBWAY_CHECK(hipGetLastError());
}

}  // namespace subway
