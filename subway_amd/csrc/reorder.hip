// reorder.hip -- renumbers a graph's vertices by descending in-degree, on
// the GPU.
//
// The relaxation kernels touch value[destination] once per edge, at random.
// On a skewed graph (R-MAT) a small set of vertices receives most of the
// edges, but in the generator's numbering those vertices are spread over the
// whole id range, so each hot 4-byte value shares its cache line with cold
// ones and the hot set occupies far more lines than the caches hold.  In
// descending in-degree order the hot vertices are contiguous at the front:
// the lines they occupy are the fewest possible and stay in L2 / the memory
// side cache.  Every result is unchanged under the renaming (BFS levels,
// distances and widths are properties of the isomorphic graph; CC starts
// from the original ids, so the smallest original id of each component
// still wins; PageRank is the same fixpoint), so the engine maps sources in
// and values out and nothing else sees the internal numbering.
//
// Steps, all on the device:
//   1. in-degrees: one atomic add per edge destination;
//   2. the vertex order: a stable radix sort of (~in-degree, id) records,
//      so equal in-degrees keep ascending ids;
//   3. perm[original] = internal, the renumbered out-degrees and their
//      exclusive scan (the new node pointers);
//   4. the edges: internal row p is original row iperm[p], in its own order,
//      with every destination renamed through perm; each workgroup fills a
//      tile of output positions, finding each position's row by a binary
//      search restricted to the tile's row range;
//   5. every host layout the graph holds (8-byte list, packed words, split
//      arrays, node pointers, degrees) overwritten in place, plus the two
//      maps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "device.hpp"
#include "radix_sort.hpp"
#include "subway/globals.hpp"
#include "subway/gpu_error_check.hpp"
#include "subway/graph.hpp"

namespace subway {

using namespace dev;

namespace {

constexpr int OB = 256;          // threads per workgroup
constexpr int OTILE = OB * 16;   // output positions per workgroup of the row kernel

// in-degree of every vertex; ends are read with a stride (2 for 8-byte
// weighted edges, 1 for destination arrays)
__global__ __launch_bounds__(OB) void in_degree_kernel(const u32 *ends, u32 stride, u64 m, u32 *indeg)
{
    const u64 step = (u64)gridDim.x * OB;
    for (u64 i = (u64)blockIdx.x * OB + threadIdx.x; i < m; i += step) atomicAdd(&indeg[ends[i * stride]], 1u);
}

__global__ __launch_bounds__(OB) void order_keys_kernel(const u32 *indeg, u64 n, u32 *key, u32 *id)
{
    const u64 step = (u64)gridDim.x * OB;
    for (u64 v = (u64)blockIdx.x * OB + threadIdx.x; v < n; v += step) {
        key[v] = 0xFFFFFFFFu - indeg[v];
        id[v] = (u32)v;
    }
}

// perm[iperm[p]] = p and the out-degree of internal vertex p
__global__ __launch_bounds__(OB) void perm_kernel(const u32 *iperm, u64 n, const u64 *ptr, u32 *perm, u32 *ndeg)
{
    const u64 step = (u64)gridDim.x * OB;
    for (u64 p = (u64)blockIdx.x * OB + threadIdx.x; p < n; p += step) {
        const u32 u = iperm[p];
        perm[u] = (u32)p;
        ndeg[p] = (u32)(ptr[u + 1] - ptr[u]);
    }
}

__global__ void set_u64_kernel(u64 *p, u64 v) { *p = v; }

// last row p in [lo, hi] with nptr[p] <= q (nptr[lo] <= q holds)
__device__ __forceinline__ u64 row_of(const u64 *nptr, u64 lo, u64 hi, u64 q)
{
    while (lo < hi) {
        const u64 mid = lo + (hi - lo + 1) / 2;
        if (nptr[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Output positions [t * OTILE, (t + 1) * OTILE): edge q of internal row p is
// edge q - nptr[p] of original row iperm[p], destination renamed.  Ends and
// weights share the stride (8-byte weighted edges: ends at word 0, weights
// at word 1 of each pair); W is the weight type (u32 in an 8-byte edge, u8
// in a split weight array); wIn == null for unweighted graphs.
template <class W>
__global__ __launch_bounds__(OB) void relabel_rows_kernel(const u64 *nptr, u64 n, u64 m, const u32 *iperm,
                                                          const u64 *ptr, const u32 *perm, const u32 *endsIn,
                                                          const W *wIn, u32 stride, u32 *endsOut, W *wOut)
{
    __shared__ u64 s_p0, s_p1;
    const u64 q0 = (u64)blockIdx.x * OTILE;
    const u64 q1 = q0 + OTILE < m ? q0 + OTILE : m;
    if (threadIdx.x == 0) {
        s_p0 = row_of(nptr, 0, n - 1, q0);
        s_p1 = row_of(nptr, s_p0, n - 1, q1 - 1);
    }
    __syncthreads();
    const u64 p0 = s_p0, p1 = s_p1;
    for (u64 q = q0 + threadIdx.x; q < q1; q += OB) {
        const u64 p = row_of(nptr, p0, p1, q);
        const u64 src = ptr[iperm[p]] + (q - nptr[p]);
        endsOut[q * stride] = perm[endsIn[src * stride]];
        if (wIn) wOut[q * stride] = wIn[src * stride];
    }
}

// derived copies of renumbered weighted edges (packed words, split arrays)
__global__ __launch_bounds__(OB) void pack_words_kernel(const uint2 *e, u64 m, u32 shift, u32 base, u32 *out)
{
    const u64 step = (u64)gridDim.x * OB;
    for (u64 i = (u64)blockIdx.x * OB + threadIdx.x; i < m; i += step) {
        const uint2 x = e[i];
        out[i] = x.x | ((x.y - base) << shift);
    }
}

__global__ __launch_bounds__(OB) void pack_split_kernel(const u32 *dst, const uint8_t *w, u64 m, u32 shift,
                                                        u32 base, u32 *out)
{
    const u64 step = (u64)gridDim.x * OB;
    for (u64 i = (u64)blockIdx.x * OB + threadIdx.x; i < m; i += step) out[i] = dst[i] | (((u32)w[i] - base) << shift);
}

__global__ __launch_bounds__(OB) void split_words_kernel(const uint2 *e, u64 m, u32 *dst, uint8_t *w)
{
    const u64 step = (u64)gridDim.x * OB;
    for (u64 i = (u64)blockIdx.x * OB + threadIdx.x; i < m; i += step) {
        const uint2 x = e[i];
        dst[i] = x.x;
        w[i] = (uint8_t)x.y;
    }
}

inline unsigned grid_of(u64 items)
{
    return (unsigned)std::max<u64>(1, std::min<u64>((items + OB - 1) / OB, 1u << 16));
}

}  // namespace

template <class E>
void reorder_by_in_degree(HostCsr<E> &csr)
{
    constexpr bool weighted = sizeof(E) == 8;
    if (csr.perm) throw Error(ERR_ARG, "the graph is already renumbered");
    const u64 n = csr.n, m = csr.m;
    const bool list = csr.edges != nullptr;
    if (m > 0 && !list && !csr.splitDst) throw Error(ERR_ARG, "the graph holds no edge list or split arrays");
    const u32 stride = list && weighted ? 2u : 1u;
    // device bytes: the edges in and out, packed / split derivations reuse the
    // input buffer; per vertex key, id (twice each), in-degree, perm, out-degree
    // and two offset arrays; the sort's tile counts
    const u64 edgeBytes = list ? m * sizeof(E) : m * (weighted ? 5 : 4);
    const u64 tiles = (n + RTILE - 1) / RTILE;
    const u64 need = 2 * edgeBytes + n * (4 * 7 + 8 * 2) + tiles * 256 * 12 + (64ull << 20);
    size_t freeB = 0, totalB = 0;
    SUBWAY_CHECK(hipMemGetInfo(&freeB, &totalB));
    if (need > (u64)freeB)
        throw Error(ERR_NOMEM, "renumbering needs " + std::to_string(need) + " bytes of device memory, " +
                                   std::to_string(freeB) + " free");
    hipStream_t st = nullptr;
    SUBWAY_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};

    DevArr<u64> dptr(n + 1), nptr(n + 1);
    SUBWAY_CHECK(hipMemcpyAsync(dptr.p, csr.ptr, (n + 1) * sizeof(u64), hipMemcpyHostToDevice, st));
    // the edges as the host holds them
    DevArr<u32> ein;
    DevArr<uint8_t> win8;
    if (list) {
        ein.alloc(m * stride);
        SUBWAY_CHECK(hipMemcpyAsync(ein.p, csr.edges, m * sizeof(E), hipMemcpyHostToDevice, st));
    } else if (m) {
        ein.alloc(m);
        SUBWAY_CHECK(hipMemcpyAsync(ein.p, csr.splitDst, m * sizeof(u32), hipMemcpyHostToDevice, st));
        if (weighted) {
            win8.alloc(m);
            SUBWAY_CHECK(hipMemcpyAsync(win8.p, csr.splitW, m, hipMemcpyHostToDevice, st));
        }
    }

    // 1-2. in-degrees, then the order
    DevArr<u32> keyA(n), idA(n), keyB(n), idB(n);
    SUBWAY_CHECK(hipMemsetAsync(keyB.p, 0, n * sizeof(u32), st));   // the in-degrees
    if (m) hipLaunchKernelGGL(in_degree_kernel, dim3(grid_of(m)), dim3(OB), 0, st, ein.p, stride, m, keyB.p);
    hipLaunchKernelGGL(order_keys_kernel, dim3(grid_of(n)), dim3(OB), 0, st, keyB.p, n, keyA.p, idA.p);
    SUBWAY_CHECK(hipGetLastError());
    u32 *key = keyA.p, *id = idA.p, *key2 = keyB.p, *id2 = idB.p;
    uint8_t *nw = nullptr, *nw2 = nullptr;
    sort_records(key, id, nw, key2, id2, nw2, n, 32, st);
    const u32 *iperm = id;
    u32 *perm = key2;    // the other key array is free now
    u32 *ndeg = key;     // and so are the sorted keys

    // 3. maps, out-degrees, node pointers
    hipLaunchKernelGGL(perm_kernel, dim3(grid_of(n)), dim3(OB), 0, st, iperm, n, dptr.p, perm, ndeg);
    SUBWAY_CHECK(hipGetLastError());
    scan_counts(ndeg, n, nptr.p, st);
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, st, nptr.p + n, (u64)m);
    SUBWAY_CHECK(hipGetLastError());

    // 4. the renumbered edges
    DevArr<u32> eout;
    DevArr<uint8_t> wout8;
    if (m) {
        eout.alloc(m * stride);
        const unsigned grid = (unsigned)((m + OTILE - 1) / OTILE);
        if (list && weighted)
            hipLaunchKernelGGL(relabel_rows_kernel<u32>, dim3(grid), dim3(OB), 0, st, nptr.p, n, m, iperm, dptr.p,
                               perm, ein.p, ein.p + 1, stride, eout.p, eout.p + 1);
        else if (weighted) {
            wout8.alloc(m);
            hipLaunchKernelGGL(relabel_rows_kernel<uint8_t>, dim3(grid), dim3(OB), 0, st, nptr.p, n, m, iperm,
                               dptr.p, perm, ein.p, win8.p, stride, eout.p, wout8.p);
        } else
            hipLaunchKernelGGL(relabel_rows_kernel<u32>, dim3(grid), dim3(OB), 0, st, nptr.p, n, m, iperm, dptr.p,
                               perm, ein.p, (const u32 *)nullptr, stride, eout.p, (u32 *)nullptr);
        SUBWAY_CHECK(hipGetLastError());
    }

    // 5. host layouts, in place
    if (m) {
        if (list) {
            SUBWAY_CHECK(hipMemcpyAsync(csr.edges, eout.p, m * sizeof(E), hipMemcpyDeviceToHost, st));
            if constexpr (weighted) {
                // the derived copies, built in the input buffer (8 bytes per edge)
                const uint2 *e2 = reinterpret_cast<const uint2 *>(eout.p);
                if (csr.packed) {
                    hipLaunchKernelGGL(pack_words_kernel, dim3(grid_of(m)), dim3(OB), 0, st, e2, m, csr.packShift,
                                       csr.packBase, ein.p);
                    SUBWAY_CHECK(hipGetLastError());
                    SUBWAY_CHECK(hipMemcpyAsync(csr.packed, ein.p, m * sizeof(u32), hipMemcpyDeviceToHost, st));
                    SUBWAY_CHECK(hipStreamSynchronize(st));
                }
                if (csr.splitDst) {
                    uint8_t *w8 = reinterpret_cast<uint8_t *>(ein.p + m);
                    hipLaunchKernelGGL(split_words_kernel, dim3(grid_of(m)), dim3(OB), 0, st, e2, m, ein.p, w8);
                    SUBWAY_CHECK(hipGetLastError());
                    SUBWAY_CHECK(hipMemcpyAsync(csr.splitDst, ein.p, m * sizeof(u32), hipMemcpyDeviceToHost, st));
                    SUBWAY_CHECK(hipMemcpyAsync(csr.splitW, w8, m, hipMemcpyDeviceToHost, st));
                }
            }
        } else {
            SUBWAY_CHECK(hipMemcpyAsync(csr.splitDst, eout.p, m * sizeof(u32), hipMemcpyDeviceToHost, st));
            if (weighted) SUBWAY_CHECK(hipMemcpyAsync(csr.splitW, wout8.p, m, hipMemcpyDeviceToHost, st));
            if (weighted && csr.packed) {
                // a loaded file's packed copy beside its split arrays, rebuilt
                // in the input buffer
                hipLaunchKernelGGL(pack_split_kernel, dim3(grid_of(m)), dim3(OB), 0, st, eout.p, wout8.p, m,
                                   csr.packShift, csr.packBase, ein.p);
                SUBWAY_CHECK(hipGetLastError());
                SUBWAY_CHECK(hipMemcpyAsync(csr.packed, ein.p, m * sizeof(u32), hipMemcpyDeviceToHost, st));
            }
        }
    }
    uint32_t *hperm = nullptr, *hiperm = nullptr;
    try {
        hperm = new uint32_t[n];
        hiperm = new uint32_t[n];
    } catch (const std::bad_alloc &) {
        delete[] hperm;
        throw Error(ERR_NOMEM, "host allocation of the vertex maps failed");
    }
    SUBWAY_CHECK(hipMemcpyAsync(csr.ptr, nptr.p, (n + 1) * sizeof(u64), hipMemcpyDeviceToHost, st));
    SUBWAY_CHECK(hipMemcpyAsync(csr.degree, ndeg, n * sizeof(u32), hipMemcpyDeviceToHost, st));
    SUBWAY_CHECK(hipMemcpyAsync(hperm, perm, n * sizeof(u32), hipMemcpyDeviceToHost, st));
    SUBWAY_CHECK(hipMemcpyAsync(hiperm, iperm, n * sizeof(u32), hipMemcpyDeviceToHost, st));
    const hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        delete[] hperm;
        delete[] hiperm;
        SUBWAY_CHECK(e);
    }
    csr.perm = hperm;
    csr.iperm = hiperm;
}

template void reorder_by_in_degree<OutEdge>(HostCsr<OutEdge> &);
template void reorder_by_in_degree<OutEdgeWeighted>(HostCsr<OutEdgeWeighted> &);

}  // namespace subway
