// kernels.hip -- gfx950 kernels of the out-of-memory graph pipeline.
//
//  * select_scan_kernel: one pass of flag -> workgroup scan -> decoupled
//    lookback -> compaction.  It implements the subgraph generator
//    (reference: prePrefix + thrust::reduce + two thrust::exclusive_scan +
//    makeQueue + makeActiveNodesPointer, shared/subgraph_generator.cu:10-60,
//    128-162) and the per-round frontier filter of every algorithm (the label
//    test/clear at the top of each kernel in shared/gpu_kernels.cu, e.g.
//    :23-28, together with mixLabels/moveUpLabels :515-534).
//  * advance_kernel: edge-balanced relaxation over the frontier's work space
//    (the edge loops of shared/gpu_kernels.cu:42-56 and siblings).
//  * partition_kernel: greedy contiguous partitioning of the active list
//    (shared/partitioner.cu:12-91).
//  * gather_kernel: zero-copy gather of active adjacency lists from mapped
//    host memory into the device partition buffer (replaces the host-thread
//    copy of shared/subgraph_generator.cu:63-97 plus the cudaMemcpy of
//    subway/sssp-async.cu:86).
#include "kernels.hpp"
#include "device.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>

#include "subway/globals.hpp"
#include "subway/gpu_error_check.hpp"

namespace subway {

using namespace dev;

// ====================================================== select-scan kernel
template <class Op, int BLOCK, int IPT>
__global__ __launch_bounds__(BLOCK) void select_scan_kernel(Op op, u64 n, ScanCtl *ctl,
                                                            TileDesc td, int first)
{
    constexpr int NW = BLOCK / WAVE;
    __shared__ u32 s_tile, s_epoch;
    __shared__ int s_exit;
    __shared__ u32 s_wcnt[NW];
    __shared__ u64 s_wsum[NW];
    __shared__ u32 s_xcnt;
    __shared__ u64 s_xsum;

    const int tid = threadIdx.x;
    if (tid == 0) {
        int ex = (!first && ld_relaxed(&ctl->stop)) ? 1 : 0;
        s_exit = ex;
        if (!ex) {
            s_epoch = ld_relaxed(&ctl->epoch);
            s_tile = fetch_add_acq_rel(&ctl->tile_counter, 1u);
        }
    }
    __syncthreads();
    if (s_exit) return;
    const u32 tile = s_tile, epoch = s_epoch;

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
        lookback(td, tile, epoch, totc, tots, xc, xs);
        if (lane == 0) {
            s_xcnt = xc;
            s_xsum = xs;
        }
    }
    __syncthreads();
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
    if (tile == gridDim.x - 1 && tid == 0) op.finish((u64)s_xcnt + totc, s_xsum + tots, ctl);
    __syncthreads();
    if (tid == 0) {
        __threadfence();
        u32 d = fetch_add_acq_rel(&ctl->done_counter, 1u);
        if (d == gridDim.x - 1) {
            st_relaxed(&ctl->done_counter, 0u);
            st_relaxed(&ctl->tile_counter, 0u);
            st_relaxed(&ctl->epoch, epoch + 1);
        }
    }
}

// ------------------------------------------------------- generator ops
struct Empty {};

// activeness = label1 || label2 (shared/subgraph_generator.cu:15)
struct GenLabelOp {
    typedef Empty Item;
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
        ctl->count = cnt;
        ctl->total = sum;
    }
};

// activeness = delta > acc (shared/subgraph_generator.cu:24-41)
struct GenPrOp {
    typedef Empty Item;
    const float *delta;
    float acc;
    const u32 *outDegree;
    u32 *activeNodes;
    u64 *activeNodesPointer;

    __device__ __forceinline__ void load(u64 i0, u64 n, Item *, u64 *w, u32 &mask) const
    {
        if (i0 + GEN_IPT <= n) {
#pragma unroll
            for (int q = 0; q < GEN_IPT / 4; q++) {
                float4 d = *reinterpret_cast<const float4 *>(delta + i0 + 4 * q);
                if (d.x > acc) mask |= 1u << (4 * q + 0);
                if (d.y > acc) mask |= 1u << (4 * q + 1);
                if (d.z > acc) mask |= 1u << (4 * q + 2);
                if (d.w > acc) mask |= 1u << (4 * q + 3);
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
        ctl->count = cnt;
        ctl->total = sum;
    }
};

// ------------------------------------------------------- frontier ops
struct FrItem {
    u32 loc;
    u32 val;
};

__device__ __forceinline__ void write_entry(Entry *entries, u32 *chunkMap, u64 k, u64 pre,
                                            u32 loc, u32 val, u64 deg)
{
    uint4 e = make_uint4(loc, (u32)pre, val, 0u);
    *reinterpret_cast<uint4 *>(entries + k) = e;
    // chunk j starts inside [pre, pre + deg): record the entry that holds it
    u64 j = (pre + ADV_CHUNK - 1) / ADV_CHUNK;
    u64 jend = (pre + deg - 1) / ADV_CHUNK;
    for (; j <= jend; j++) chunkMap[j] = (u32)k;
}

// Labelled vertices of the partition push their current value; their labels
// are cleared as the reference does before its edge loop.
struct FrontierLabelOp {
    typedef FrItem Item;
    const u32 *nodes;
    const u64 *ptr;
    u64 fromEdge;
    uint8_t *labA, *labB;
    const u32 *value;
    Entry *entries;
    u32 *chunkMap;

    __device__ __forceinline__ void load(u64 i0, u64 n, Item *it, u64 *w, u32 &mask) const
    {
#pragma unroll
        for (int j = 0; j < FR_IPT; j++) {
            u64 i = i0 + j;
            if (i < n) {
                u32 id = nodes[i];
                uint8_t f = labA[id];
                if (labB) f |= labB[id];
                if (f) {
                    labA[id] = 0;
                    if (labB) labB[id] = 0;
                    u64 p0 = ptr[i], p1 = ptr[i + 1];
                    if (p1 > p0) {
                        mask |= 1u << j;
                        w[j] = p1 - p0;
                        it[j].loc = (u32)(p0 - fromEdge);
                        it[j].val = value[id];
                    }
                }
            }
        }
    }
    __device__ __forceinline__ void emit(u64 k, u64 pre, u64, const Item &it, u64 deg) const
    {
        write_entry(entries, chunkMap, k, pre, it.loc, it.val, deg);
    }
    __device__ __forceinline__ void finish(u64 cnt, u64 sum, ScanCtl *ctl) const
    {
        ctl->count = cnt;
        ctl->total = sum;
        ctl->stop = (cnt == 0) ? 1u : 0u;
        if (cnt) ctl->rounds += 1;
        ctl->entry_sum += cnt;
        ctl->work_sum += sum;
    }
};

// PageRank: delta > acc selects; value absorbs delta and the share is
// (float)((d / deg) * 0.85) as in shared/gpu_kernels.cu:233-239.
struct FrontierPrOp {
    typedef FrItem Item;
    const u32 *nodes;
    const u64 *ptr;
    u64 fromEdge;
    float *value;
    float *delta;
    float acc;
    Entry *entries;
    u32 *chunkMap;

    __device__ __forceinline__ void load(u64 i0, u64 n, Item *it, u64 *w, u32 &mask) const
    {
#pragma unroll
        for (int j = 0; j < FR_IPT; j++) {
            u64 i = i0 + j;
            if (i < n) {
                u32 id = nodes[i];
                float d = delta[id];
                if (d > acc) {
                    value[id] += d;
                    delta[id] = 0.0f;
                    u64 p0 = ptr[i], p1 = ptr[i + 1];
                    if (p1 > p0) {
                        float share = (float)((double)(d / (float)(p1 - p0)) * 0.85);
                        mask |= 1u << j;
                        w[j] = p1 - p0;
                        it[j].loc = (u32)(p0 - fromEdge);
                        it[j].val = __float_as_uint(share);
                    }
                }
            }
        }
    }
    __device__ __forceinline__ void emit(u64 k, u64 pre, u64, const Item &it, u64 deg) const
    {
        write_entry(entries, chunkMap, k, pre, it.loc, it.val, deg);
    }
    __device__ __forceinline__ void finish(u64 cnt, u64 sum, ScanCtl *ctl) const
    {
        ctl->count = cnt;
        ctl->total = sum;
        ctl->stop = (cnt == 0) ? 1u : 0u;
        if (cnt) ctl->rounds += 1;
        ctl->entry_sum += cnt;
        ctl->work_sum += sum;
    }
};

// ============================================================ advance
template <int ALGO>
struct Relax;

// cand = value[u] + 1, keep min   (shared/gpu_kernels.cu:42-56)
template <>
struct Relax<ALGO_BFS> {
    template <class E>
    __device__ __forceinline__ static bool apply(u32 val, const E &e, u32 *value, uint8_t *lab, float *)
    {
        u32 cand = val + 1;
        u32 v = e.end;
        if (cand < value[v]) {
            u32 prev = atomicMin(&value[v], cand);
            if (cand < prev) {
                lab[v] = 1;
                return true;
            }
        }
        return false;
    }
};
// cand = value[u], keep min       (shared/gpu_kernels.cu:90-106)
template <>
struct Relax<ALGO_CC> {
    template <class E>
    __device__ __forceinline__ static bool apply(u32 val, const E &e, u32 *value, uint8_t *lab, float *)
    {
        u32 v = e.end;
        if (val < value[v]) {
            u32 prev = atomicMin(&value[v], val);
            if (val < prev) {
                lab[v] = 1;
                return true;
            }
        }
        return false;
    }
};
// cand = value[u] + w, keep min   (shared/gpu_kernels.cu:143-157)
template <>
struct Relax<ALGO_SSSP> {
    __device__ __forceinline__ static bool apply(u32 val, const OutEdgeWeighted &e, u32 *value,
                                                 uint8_t *lab, float *)
    {
        u32 cand = val + e.w8;
        u32 v = e.end;
        if (cand < value[v]) {
            u32 prev = atomicMin(&value[v], cand);
            if (cand < prev) {
                lab[v] = 1;
                return true;
            }
        }
        return false;
    }
};
// cand = min(value[u], w), keep max (shared/gpu_kernels.cu:193-207)
template <>
struct Relax<ALGO_SSWP> {
    __device__ __forceinline__ static bool apply(u32 val, const OutEdgeWeighted &e, u32 *value,
                                                 uint8_t *lab, float *)
    {
        u32 cand = val < e.w8 ? val : e.w8;
        u32 v = e.end;
        if (cand > value[v]) {
            u32 prev = atomicMax(&value[v], cand);
            if (cand > prev) {
                lab[v] = 1;
                return true;
            }
        }
        return false;
    }
};
// delta[v] += share             (shared/gpu_kernels.cu:241-244)
template <>
struct Relax<ALGO_PR> {
    template <class E>
    __device__ __forceinline__ static bool apply(u32 val, const E &e, u32 *, uint8_t *, float *delta)
    {
        unsafeAtomicAdd(&delta[e.end], __uint_as_float(val));
        return true;
    }
};

template <int ALGO, class E>
__global__ __launch_bounds__(ADV_BLOCK) void advance_kernel(const Entry *entries, const u32 *chunkMap,
                                                            const E *edges, u32 *value, uint8_t *lab,
                                                            float *delta, u64 w0, u64 w1, u64 shift,
                                                            ScanCtl *ctl)
{
    __shared__ u32 s_pre[ADV_CHUNK + 1];
    __shared__ u32 s_loc[ADV_CHUNK + 1];
    __shared__ u32 s_val[ADV_CHUNK + 1];
    __shared__ u32 s_chunk;

    const int tid = threadIdx.x;
    const u64 count = ctl->count;
    const u64 W = ctl->total;
    const u64 wend = W < w1 ? W : w1;
    if (count == 0 || wend <= w0) return;
    const u64 j0 = w0 / ADV_CHUNK;
    const u64 nchunks = (wend + ADV_CHUNK - 1) / ADV_CHUNK - j0;
    u32 updates = 0;

    for (;;) {
        if (tid == 0) s_chunk = atomicAdd(&ctl->adv_counter, 1u);
        __syncthreads();
        const u64 c = s_chunk;
        if (c >= nchunks) break;
        const u64 j = j0 + c;
        const u64 cbeg = j * ADV_CHUNK;
        const u64 pbeg = cbeg > w0 ? cbeg : w0;
        const u64 pend = (cbeg + ADV_CHUNK) < wend ? (cbeg + ADV_CHUNK) : wend;
        const u64 k0 = chunkMap[j];
        const u64 k1 = (cbeg + ADV_CHUNK < W) ? chunkMap[j + 1] : count - 1;
        const int nk = (int)(k1 - k0 + 1);
        for (int t = tid; t < nk; t += ADV_BLOCK) {
            uint4 e = *reinterpret_cast<const uint4 *>(entries + k0 + t);
            s_loc[t] = e.x;
            s_pre[t] = e.y;
            s_val[t] = e.z;
        }
        __syncthreads();
        for (u64 p = pbeg + tid; p < pend; p += ADV_BLOCK) {
            const u32 pp = (u32)p;
            int lo = 0, hi = nk - 1;
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (s_pre[mid] <= pp) lo = mid;
                else hi = mid - 1;
            }
            const u64 eidx = (u64)s_loc[lo] + (pp - s_pre[lo]) - shift;
            const E e = edges[eidx];
            if (Relax<ALGO>::apply(s_val[lo], e, value, lab, delta)) updates++;
        }
        __syncthreads();
    }
    u32 wsum = wave_reduce(updates);
    if (lane_id() == 0 && wsum) atomicAdd((unsigned long long *)&ctl->updates, (unsigned long long)wsum);
    if (tid == 0) {
        __threadfence();
        u32 d = fetch_add_acq_rel(&ctl->adv_done, 1u);
        if (d == gridDim.x - 1) {
            st_relaxed(&ctl->adv_done, 0u);
            st_relaxed(&ctl->adv_counter, 0u);
        }
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
    const u64 A = dA ? *dA : hostA;
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

// ============================================================== gather
// Destination positions [q0, q1) of the partition are mapped to the active
// vertex that owns them through LDS-staged segment starts; sources are read
// straight from mapped pinned host memory.
constexpr int GATHER_SEGS = 1024;

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
    const u64 q0 = (u64)blockIdx.x * GATHER_CHUNK;
    if (q0 >= edgeCount) return;
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
    SUBWAY_CHECK(hipMalloc(&s.td.agg, maxTiles * sizeof(u64)));
    SUBWAY_CHECK(hipMalloc(&s.td.inc, maxTiles * sizeof(u64)));
    // epoch tags start at 0 in the control block; descriptors are cleared to
    // an impossible tag so stale memory never reads as published
    SUBWAY_CHECK(hipMemset(s.td.flag, 0xFF, maxTiles * sizeof(u64)));
    SUBWAY_CHECK(hipDeviceSynchronize());
    s.maxTiles = maxTiles;
}

void free_scan_scratch(ScanScratch &s)
{
    if (s.ctl) (void)hipFree(s.ctl);
    if (s.td.flag) (void)hipFree(s.td.flag);
    if (s.td.agg) (void)hipFree(s.td.agg);
    if (s.td.inc) (void)hipFree(s.td.inc);
    s = ScanScratch{};
}

static void check_tiles(const ScanScratch &s, u64 tiles)
{
    if (tiles > s.maxTiles)
        throw Error(ERR_INTERNAL, "scan launch needs " + std::to_string(tiles) +
                                      " tiles but scratch holds " + std::to_string(s.maxTiles));
    if (tiles > 0x7FFFFFFFull) throw Error(ERR_ARG, "scan launch too large");
}

void launch_generate_labels(ull n, const uint8_t *l1, const uint8_t *l2,
                            const uint32_t *outDegree, uint32_t *activeNodes,
                            ull *activeNodesPointer, ScanScratch &s, hipStream_t st)
{
    if (n == 0) throw Error(ERR_ARG, "generate on an empty graph");
    u64 tiles = (n + GEN_TILE - 1) / GEN_TILE;
    check_tiles(s, tiles);
    GenLabelOp op{l1, l2, outDegree, activeNodes, activeNodesPointer};
    hipLaunchKernelGGL((select_scan_kernel<GenLabelOp, SCAN_BLOCK, GEN_IPT>), dim3((unsigned)tiles),
                       dim3(SCAN_BLOCK), 0, st, op, (u64)n, s.ctl, s.td, 1);
    SUBWAY_CHECK(hipGetLastError());
}

void launch_generate_pr(ull n, const float *delta, float acc, const uint32_t *outDegree,
                        uint32_t *activeNodes, ull *activeNodesPointer, ScanScratch &s,
                        hipStream_t st)
{
    if (n == 0) throw Error(ERR_ARG, "generate on an empty graph");
    u64 tiles = (n + GEN_TILE - 1) / GEN_TILE;
    check_tiles(s, tiles);
    GenPrOp op{delta, acc, outDegree, activeNodes, activeNodesPointer};
    hipLaunchKernelGGL((select_scan_kernel<GenPrOp, SCAN_BLOCK, GEN_IPT>), dim3((unsigned)tiles),
                       dim3(SCAN_BLOCK), 0, st, op, (u64)n, s.ctl, s.td, 1);
    SUBWAY_CHECK(hipGetLastError());
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
    u64 blocks = (edgeCount + GATHER_CHUNK - 1) / GATHER_CHUNK;
    if (blocks > 0x7FFFFFFFull) throw Error(ERR_ARG, "gather too large");
    if (edgeBytes == 8)
        hipLaunchKernelGGL((gather_kernel<OutEdgeWeighted>), dim3((unsigned)blocks), dim3(GATHER_BLOCK), 0,
                           st, activeNodes, activeNodesPointer, (u64)kbeg, (u64)kend, nodePointer,
                           (const OutEdgeWeighted *)hostEdges, (OutEdgeWeighted *)dst, (u64)edgeBase,
                           (u64)edgeCount);
    else
        hipLaunchKernelGGL((gather_kernel<OutEdge>), dim3((unsigned)blocks), dim3(GATHER_BLOCK), 0, st,
                           activeNodes, activeNodesPointer, (u64)kbeg, (u64)kend, nodePointer,
                           (const OutEdge *)hostEdges, (OutEdge *)dst, (u64)edgeBase, (u64)edgeCount);
    SUBWAY_CHECK(hipGetLastError());
}

void launch_frontier(const FrontierArgs &a, ScanScratch &s, hipStream_t st)
{
    if (a.n == 0) throw Error(ERR_ARG, "frontier over an empty partition");
    u64 tiles = (a.n + FR_TILE - 1) / FR_TILE;
    check_tiles(s, tiles);
    if (a.algo == ALGO_PR) {
        FrontierPrOp op{a.nodes, a.ptr, a.fromEdge, (float *)a.value, a.delta, a.acc, a.entries, a.chunkMap};
        hipLaunchKernelGGL((select_scan_kernel<FrontierPrOp, SCAN_BLOCK, FR_IPT>), dim3((unsigned)tiles),
                           dim3(SCAN_BLOCK), 0, st, op, (u64)a.n, s.ctl, s.td, a.first);
    } else {
        FrontierLabelOp op{a.nodes, a.ptr, a.fromEdge, a.labA, a.labB, a.value, a.entries, a.chunkMap};
        hipLaunchKernelGGL((select_scan_kernel<FrontierLabelOp, SCAN_BLOCK, FR_IPT>),
                           dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, st, op, (u64)a.n, s.ctl, s.td,
                           a.first);
    }
    SUBWAY_CHECK(hipGetLastError());
}

template <int ALGO, class E>
static void adv(const AdvanceArgs &a, ScanCtl *ctl, hipStream_t st)
{
    hipLaunchKernelGGL((advance_kernel<ALGO, E>), dim3((unsigned)a.grid), dim3(ADV_BLOCK), 0, st,
                       a.entries, a.chunkMap, (const E *)a.edges, a.value, a.labelOut, a.delta,
                       (u64)a.w0, (u64)a.w1, (u64)a.shift, ctl);
}

void launch_advance(const AdvanceArgs &a, ScanCtl *ctl, hipStream_t st)
{
    if (a.grid <= 0) throw Error(ERR_ARG, "advance grid must be positive");
    switch (a.algo) {
        case ALGO_BFS:
            if (a.edgeBytes == 4) adv<ALGO_BFS, OutEdge>(a, ctl, st);
            else adv<ALGO_BFS, OutEdgeWeighted>(a, ctl, st);
            break;
        case ALGO_CC:
            if (a.edgeBytes == 4) adv<ALGO_CC, OutEdge>(a, ctl, st);
            else adv<ALGO_CC, OutEdgeWeighted>(a, ctl, st);
            break;
        case ALGO_SSSP:
            if (a.edgeBytes != 8) throw Error(ERR_ARG, "SSSP needs weighted edges");
            adv<ALGO_SSSP, OutEdgeWeighted>(a, ctl, st);
            break;
        case ALGO_SSWP:
            if (a.edgeBytes != 8) throw Error(ERR_ARG, "SSWP needs weighted edges");
            adv<ALGO_SSWP, OutEdgeWeighted>(a, ctl, st);
            break;
        case ALGO_PR:
            if (a.edgeBytes == 4) adv<ALGO_PR, OutEdge>(a, ctl, st);
            else adv<ALGO_PR, OutEdgeWeighted>(a, ctl, st);
            break;
        default: throw Error(ERR_ARG, "unknown algorithm");
    }
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
