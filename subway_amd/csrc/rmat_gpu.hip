// rmat_gpu.hip -- the R-MAT generator of rmat.cpp, built on the GPU.
//
// The host generator takes minutes at 2^33 edges (two hash passes over
// every edge, a 12-byte-per-edge bucket array, a counting sort); this one
// builds the byte-identical CSR on the device and copies it into the pinned
// host arrays:
//   1. records: edge e (and, symmetrised, its reverse) becomes a record
//      (key = source, destination, weight) at index e * copies + copy, in
//      generation order; a self-loop gets key n, past every vertex;
//   2. a stable LSD radix sort of the records by key (8-bit digits at most;
//      per digit pass: tile histograms, an exclusive scan of them in
//      digit-major order, a stable scatter whose in-tile ranks come from
//      wave64 ballots), so every vertex's records stay in generation order
//      -- the order of the host generator's stable counting sort;
//   3. node pointers from the key boundaries of the sorted records; the
//      self-loops sit past the last vertex and are dropped;
//   4. the edge arrays in the graph's host layout (8-byte edges, destinations
//      only, or the split destination + weight arrays) copied to the host.
// The hash functions are the host generator's, so every (source,
// destination, weight) triple is the same.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "device.hpp"
#include "radix_sort.hpp"
#include "host_util.hpp"
#include "subway/globals.hpp"
#include "subway/gpu_error_check.hpp"
#include "subway/graph.hpp"

namespace subway {

using namespace dev;

namespace {

struct RmatDev {
    int scale;
    u64 nominal;          // edgeFactor * 2^scale
    u32 ta, tab, tabc;    // cumulative quadrant thresholds on 2^32
    u64 seed;
    u32 wmin, wspan;
    u32 copies;           // 2 when symmetrised
    u32 weighted;
};

__device__ __forceinline__ u64 mix64(u64 z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// The host generator's edge function (rmat.cpp): one 64-bit hash per two
// levels, 32 bits per level against the quadrant thresholds.
__device__ __forceinline__ bool rmat_edge(const RmatDev &p, u64 e, u32 &src, u32 &dst)
{
    u32 s = 0, d = 0;
    const u64 key = mix64(p.seed ^ 0x5DEECE66Dull);
    for (int l = 0; l < p.scale; l += 2) {
        const u64 h = mix64(key ^ mix64((e << 5) | (u64)(l >> 1)));
        for (int k = 0; k < 2 && l + k < p.scale; k++) {
            const u32 r = (u32)(h >> (32 * k));
            const u32 bit = 1u << (p.scale - 1 - (l + k));
            if (r < p.ta) {
            } else if (r < p.tab) {
                d |= bit;
            } else if (r < p.tabc) {
                s |= bit;
            } else {
                s |= bit;
                d |= bit;
            }
        }
    }
    src = s;
    dst = d;
    return s != d;
}

__device__ __forceinline__ u32 rmat_weight(const RmatDev &p, u64 e)
{
    const u64 h = mix64(mix64(p.seed ^ 0xC2B2AE3D27D4EB4Full) ^ e);
    return p.wmin + (u32)(h % p.wspan);
}

__global__ __launch_bounds__(RB) void rmat_records_kernel(RmatDev p, u32 n, u32 *key, u32 *dst, uint8_t *w)
{
    const u64 stride = (u64)gridDim.x * RB;
    for (u64 e = (u64)blockIdx.x * RB + threadIdx.x; e < p.nominal; e += stride) {
        u32 s, d;
        const bool live = rmat_edge(p, e, s, d);
        const u32 wt = p.weighted ? rmat_weight(p, e) : 1u;
        const u64 r = e * p.copies;
        key[r] = live ? s : n;
        dst[r] = d;
        if (w) w[r] = (uint8_t)wt;
        if (p.copies == 2) {
            key[r + 1] = live ? d : n;
            dst[r + 1] = s;
            if (w) w[r + 1] = (uint8_t)wt;
        }
    }
}

// Node pointers from the sorted keys: ptr[v] is the first record whose key
// is v or more, so record i starts every vertex in (key[i-1], key[i]] (all
// vertices up to key[0] for i = 0, those after the last key up to n for
// i = R).  Self-loops carry key n: ptr[n], the edge count, is the first of
// them (or R).  Every vertex is written exactly once.
__global__ __launch_bounds__(RB) void rmat_ptr_kernel(const u32 *key, u64 R, u32 n, u64 *ptr)
{
    const u64 stride = (u64)gridDim.x * RB;
    for (u64 i = (u64)blockIdx.x * RB + threadIdx.x; i <= R; i += stride) {
        const u64 lo = i == 0 ? 0ull : (u64)key[i - 1] + 1ull;
        const u64 hi = i == R ? (u64)n : (u64)key[i];
        for (u64 v = lo; v <= hi && v <= (u64)n; v++) ptr[v] = i;
    }
}

__global__ __launch_bounds__(RB) void interleave_kernel(const u32 *dst, const uint8_t *w, u64 m, uint2 *out)
{
    const u64 stride = (u64)gridDim.x * RB;
    for (u64 i = (u64)blockIdx.x * RB + threadIdx.x; i < m; i += stride) out[i] = make_uint2(dst[i], (u32)w[i]);
}

}  // namespace

// Largest R-MAT the device can build: records of 9 bytes (5 unweighted)
// twice over, plus the per-tile digit counts and offsets.
static u64 rmat_gpu_bytes(u64 R, bool weighted)
{
    const u64 per = weighted ? 9 : 8;
    const u64 tiles = (R + RTILE - 1) / RTILE;
    return 2 * per * R + tiles * 256 * 12 + (64ull << 20);
}

template <class E>
std::shared_ptr<HostCsr<E>> rmat_csr_gpu(int scale, int edgeFactor, double a, double b, double c, uint64_t seed,
                                         bool symmetrize, uint32_t wmin, uint32_t wmax, bool splitLayout)
{
    if (scale < 1 || scale > 31) throw Error(ERR_ARG, "R-MAT scale must be in [1, 31]");
    if (edgeFactor < 1) throw Error(ERR_ARG, "R-MAT edge factor must be positive");
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) throw Error(ERR_ARG, "R-MAT probabilities invalid");
    if (wmin < 1 || wmax < wmin) throw Error(ERR_ARG, "R-MAT weight range invalid");
    constexpr bool weighted = sizeof(E) == 8;
    if (weighted && wmax > 0xFFu)
        throw Error(ERR_ARG, "the GPU R-MAT build keeps weights in a byte (wmax <= 255); build on the host");
    RmatDev p{};
    p.scale = scale;
    p.nominal = (u64)edgeFactor << scale;
    auto thr = [](double x) -> u32 {
        const double v = x * 4294967296.0;
        return v >= 4294967295.0 ? 0xFFFFFFFFu : (u32)v;
    };
    p.ta = thr(a);
    p.tab = thr(a + b);
    p.tabc = thr(a + b + c);
    p.seed = seed;
    p.wmin = wmin;
    p.wspan = wmax - wmin + 1;
    p.copies = symmetrize ? 2u : 1u;
    p.weighted = weighted ? 1u : 0u;
    const u64 n = 1ull << scale;
    if (n > 0xFFFFFFFFull) throw Error(ERR_ARG, "R-MAT scale too large for 32-bit ids");
    const u64 R = p.nominal * p.copies;
    size_t freeB = 0, totalB = 0;
    SUBWAY_CHECK(hipMemGetInfo(&freeB, &totalB));
    if (rmat_gpu_bytes(R, weighted) > (u64)freeB)
        throw Error(ERR_NOMEM, "GPU R-MAT build needs " + std::to_string(rmat_gpu_bytes(R, weighted)) +
                                   " bytes of device memory, " + std::to_string(freeB) + " free");
    hipStream_t st = nullptr;
    SUBWAY_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};

    // 1. records in generation order
    DevArr<u32> keyA(R), dstA(R), keyB(R), dstB(R);
    DevArr<uint8_t> wA, wB;
    if (weighted) {
        wA.alloc(R);
        wB.alloc(R);
    }
    hipLaunchKernelGGL(rmat_records_kernel, dim3(grid_for(p.nominal, 1u << 16)), dim3(RB), 0, st, p, (u32)n,
                       keyA.p, dstA.p, wA.p);
    SUBWAY_CHECK(hipGetLastError());

    // 2. stable LSD radix sort by key over scale + 1 bits (self-loops carry key n)
    u32 *kin = keyA.p, *din = dstA.p, *kout = keyB.p, *dout = dstB.p;
    uint8_t *win = wA.p, *wout = wB.p;
    sort_records(kin, din, win, kout, dout, wout, R, scale + 1, st);

    // free the scratch pair
    const bool inA = kin == keyA.p;
    if (inA) {
        keyB.reset();
        dstB.reset();
        wB.reset();
    } else {
        keyA.reset();
        dstA.reset();
        wA.reset();
    }

    // 3. node pointers and the edge count
    DevArr<u64> dptr(n + 1);
    hipLaunchKernelGGL(rmat_ptr_kernel, dim3(grid_for(R + 1)), dim3(RB), 0, st, kin, R, (u32)n, dptr.p);
    SUBWAY_CHECK(hipGetLastError());
    u64 m = 0;
    SUBWAY_CHECK(hipMemcpyAsync(&m, dptr.p + n, sizeof(u64), hipMemcpyDeviceToHost, st));
    SUBWAY_CHECK(hipStreamSynchronize(st));
    if (m > R) throw Error(ERR_INTERNAL, "GPU R-MAT build: edge count out of range");
    (inA ? keyA : keyB).reset();

    // 4. host layout, as the host generator lays it out
    auto csr = std::make_shared<HostCsr<E>>();
    bool split = false;
    if constexpr (weighted) {
        const bool packs = (u64)(wmax - wmin) < (1ull << (32 - std::min(scale, 31)));
        split = splitLayout || (!packs && (u64)m * sizeof(E) >= (1ull << 34));
    }
    // a split-layout graph built here keeps no 8-byte list on the host: its
    // edges are the pinned destination and weight arrays
    csr->allocate(n, m, !split, /*withEdges=*/!split);
    if (split && !csr->allocateSplit()) throw Error(ERR_NOMEM, "pinned allocation of the split edge arrays failed");
    csr->large = m > 0xFFFFFFFFull;
    SUBWAY_CHECK(hipMemcpyAsync(csr->ptr, dptr.p, (n + 1) * sizeof(u64), hipMemcpyDeviceToHost, st));
    if (split) {
        SUBWAY_CHECK(hipMemcpyAsync(csr->splitDst, din, m * sizeof(u32), hipMemcpyDeviceToHost, st));
        SUBWAY_CHECK(hipMemcpyAsync(csr->splitW, win, m, hipMemcpyDeviceToHost, st));
    } else if constexpr (weighted) {
        DevArr<uint2> out(m);
        hipLaunchKernelGGL(interleave_kernel, dim3(grid_for(m)), dim3(RB), 0, st, din, win, m, out.p);
        SUBWAY_CHECK(hipGetLastError());
        SUBWAY_CHECK(hipMemcpyAsync(csr->edges, out.p, m * sizeof(E), hipMemcpyDeviceToHost, st));
        SUBWAY_CHECK(hipStreamSynchronize(st));
    } else {
        SUBWAY_CHECK(hipMemcpyAsync(csr->edges, din, m * sizeof(E), hipMemcpyDeviceToHost, st));
    }
    SUBWAY_CHECK(hipStreamSynchronize(st));
    csr->finalize();
    return csr;
}

template std::shared_ptr<HostCsr<OutEdge>> rmat_csr_gpu<OutEdge>(int, int, double, double, double, uint64_t, bool,
                                                                 uint32_t, uint32_t, bool);
template std::shared_ptr<HostCsr<OutEdgeWeighted>> rmat_csr_gpu<OutEdgeWeighted>(int, int, double, double, double,
                                                                                 uint64_t, bool, uint32_t, uint32_t,
                                                                                 bool);

}  // namespace subway
