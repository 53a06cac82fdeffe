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

constexpr int RB = 256;                 // threads per workgroup of the sort kernels
constexpr int RIPT = 32;                // records per thread and tile
constexpr int RTILE = RB * RIPT;        // records per tile
constexpr int NWAVE = RB / WAVE;
constexpr int SCAN_IPT = 16;
constexpr int SCAN_TILE = RB * SCAN_IPT;

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

// Histogram of one digit per tile, stored digit-major: counts[b * tiles + t].
__global__ __launch_bounds__(RB) void radix_hist_kernel(const u32 *key, u64 R, int shift, int bins, u32 *counts,
                                                        u64 tiles)
{
    __shared__ u32 h[256];
    const u64 t = blockIdx.x;
    for (int b = threadIdx.x; b < 256; b += RB) h[b] = 0;
    __syncthreads();
    const u64 i0 = t * RTILE;
    const u32 mask = (u32)bins - 1u;
#pragma unroll 4
    for (int q = 0; q < RIPT; q++) {
        const u64 i = i0 + (u64)q * RB + threadIdx.x;
        if (i < R) atomicAdd(&h[(key[i] >> shift) & mask], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < bins; b += RB) counts[(u64)b * tiles + t] = h[b];
}

// Exclusive scan of u32 counts into u64 offsets, in three launches: block
// sums, a one-workgroup scan of the block sums, and the block scans.
__global__ __launch_bounds__(RB) void scan_sum_kernel(const u32 *in, u64 N, u64 *part)
{
    __shared__ u64 s[NWAVE];
    const u64 i0 = (u64)blockIdx.x * SCAN_TILE + (u64)threadIdx.x * SCAN_IPT;
    u64 v = 0;
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++)
        if (i0 + j < N) v += in[i0 + j];
    v = wave_reduce(v);
    if (lane_id() == 0) s[threadIdx.x / WAVE] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 t = 0;
        for (int q = 0; q < NWAVE; q++) t += s[q];
        part[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(1024) void scan_parts_kernel(u64 *part, u64 P, u64 *total)
{
    constexpr int T = 1024, NW = T / WAVE;
    __shared__ u64 s[NW];
    const int tid = threadIdx.x, wid = tid / WAVE, lane = lane_id();
    const u64 per = (P + T - 1) / T;
    const u64 b0 = (u64)tid * per < P ? (u64)tid * per : P;
    const u64 b1 = b0 + per < P ? b0 + per : P;
    u64 c = 0;
    for (u64 b = b0; b < b1; b++) c += part[b];
    const u64 ic = wave_incl_scan(c);
    if (lane == WAVE - 1) s[wid] = ic;
    __syncthreads();
    u64 bc = 0, tc = 0;
    for (int q = 0; q < NW; q++) {
        if (q < wid) bc += s[q];
        tc += s[q];
    }
    u64 x = bc + ic - c;
    for (u64 b = b0; b < b1; b++) {
        const u64 v = part[b];
        part[b] = x;
        x += v;
    }
    if (tid == 0) *total = tc;
}

__global__ __launch_bounds__(RB) void scan_apply_kernel(const u32 *in, u64 N, const u64 *part, u64 *out)
{
    __shared__ u64 s[NWAVE];
    const u64 i0 = (u64)blockIdx.x * SCAN_TILE + (u64)threadIdx.x * SCAN_IPT;
    u32 v[SCAN_IPT];
    u64 sum = 0;
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++) {
        v[j] = i0 + j < N ? in[i0 + j] : 0u;
        sum += v[j];
    }
    const u64 inc = wave_incl_scan(sum);
    const int wid = threadIdx.x / WAVE;
    if (lane_id() == WAVE - 1) s[wid] = inc;
    __syncthreads();
    u64 x = part[blockIdx.x] + inc - sum;
    for (int q = 0; q < wid; q++) x += s[q];
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++) {
        if (i0 + j < N) out[i0 + j] = x;
        x += v[j];
    }
}

// Stable scatter of one digit pass.  A tile's records are taken in rounds of
// RB consecutive records (one per thread, in thread order); within a round a
// record's rank among equal digits is the count of equal-digit lanes before
// it in its wave (ballot match over the digit's bits) plus the equal-digit
// counts of the waves before it; rounds follow one another, so the output
// keeps the input order among equal digits.
__global__ __launch_bounds__(RB) void radix_scatter_kernel(const u32 *keyIn, const u32 *dstIn, const uint8_t *wIn,
                                                           u32 *keyOut, u32 *dstOut, uint8_t *wOut, u64 R,
                                                           int shift, int dbits, const u64 *offs, u64 tiles)
{
    __shared__ u64 s_base[256];
    __shared__ u32 s_wcnt[NWAVE][256];
    const int tid = threadIdx.x, wid = tid / WAVE, lane = lane_id();
    const int bins = 1 << dbits;
    const u32 mask = (u32)bins - 1u;
    const u64 t = blockIdx.x;
    for (int b = tid; b < 256; b += RB) {
        s_base[b] = b < bins ? offs[(u64)b * tiles + t] : 0ull;
#pragma unroll
        for (int q = 0; q < NWAVE; q++) s_wcnt[q][b] = 0;
    }
    __syncthreads();
    const u64 lanesBelow = (1ull << lane) - 1ull;
    const u64 i0 = t * RTILE;
    for (int q = 0; q < RIPT; q++) {
        const u64 base = i0 + (u64)q * RB;
        if (base >= R) break;   // uniform across the workgroup
        const u64 i = base + tid;
        const bool valid = i < R;
        u32 k = 0, d = 0;
        uint8_t wt = 0;
        if (valid) {
            k = keyIn[i];
            d = dstIn[i];
            if (wIn) wt = wIn[i];
        }
        const u32 dig = (k >> shift) & mask;
        u64 peers = __ballot(valid);
        for (int b = 0; b < dbits; b++) {
            const u64 set = __ballot(valid && ((dig >> b) & 1u));
            peers &= ((dig >> b) & 1u) ? set : ~set;
        }
        const u32 rank = (u32)__popcll(peers & lanesBelow);
        if (valid && rank == 0) s_wcnt[wid][dig] = (u32)__popcll(peers);
        __syncthreads();
        if (valid) {
            u64 pos = s_base[dig] + rank;
            for (int w2 = 0; w2 < wid; w2++) pos += s_wcnt[w2][dig];
            keyOut[pos] = k;
            dstOut[pos] = d;
            if (wOut) wOut[pos] = wt;
        }
        __syncthreads();
        for (int b = tid; b < bins; b += RB) {
            u64 add = 0;
#pragma unroll
            for (int w2 = 0; w2 < NWAVE; w2++) {
                add += s_wcnt[w2][b];
                s_wcnt[w2][b] = 0;
            }
            s_base[b] += add;
        }
        __syncthreads();
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

template <class T>
struct DevArr {
    T *p = nullptr;
    DevArr() = default;
    explicit DevArr(u64 n) { alloc(n); }
    void alloc(u64 n)
    {
        const hipError_t e = hipMalloc(&p, (size_t)(n ? n : 1) * sizeof(T));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            throw Error(ERR_NOMEM, "GPU R-MAT build: device allocation of " + std::to_string(n * sizeof(T)) +
                                       " bytes failed (" + hipGetErrorString(e) + ")");
        }
    }
    void reset()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    ~DevArr() { reset(); }
    DevArr(const DevArr &) = delete;
    DevArr &operator=(const DevArr &) = delete;
};

inline unsigned grid_for(u64 items, unsigned cap = 65536)
{
    const u64 g = (items + RB - 1) / RB;
    return (unsigned)std::max<u64>(1, std::min<u64>(g, cap));
}

// Exclusive scan of N u32 counts into u64 offsets (out may not alias in).
void scan_counts(const u32 *in, u64 N, u64 *out, hipStream_t st)
{
    const u64 P = (N + SCAN_TILE - 1) / SCAN_TILE;
    DevArr<u64> part(P + 1);
    hipLaunchKernelGGL(scan_sum_kernel, dim3((unsigned)P), dim3(RB), 0, st, in, N, part.p);
    hipLaunchKernelGGL(scan_parts_kernel, dim3(1), dim3(1024), 0, st, part.p, P, part.p + P);
    hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)P), dim3(RB), 0, st, in, N, part.p, out);
    SUBWAY_CHECK(hipGetLastError());
    SUBWAY_CHECK(hipStreamSynchronize(st));
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
    const int keyBits = scale + 1;
    const int passes = (keyBits + 7) / 8;
    const u64 tiles = (R + RTILE - 1) / RTILE;
    DevArr<u32> counts(tiles * 256);
    DevArr<u64> offs(tiles * 256);
    u32 *kin = keyA.p, *din = dstA.p, *kout = keyB.p, *dout = dstB.p;
    uint8_t *win = wA.p, *wout = wB.p;
    int shift = 0;
    for (int ps = 0; ps < passes; ps++) {
        const int db = (keyBits - shift + (passes - ps) - 1) / (passes - ps);   // spread the bits evenly
        const int bins = 1 << db;
        hipLaunchKernelGGL(radix_hist_kernel, dim3((unsigned)tiles), dim3(RB), 0, st, kin, R, shift, bins,
                           counts.p, tiles);
        SUBWAY_CHECK(hipGetLastError());
        scan_counts(counts.p, (u64)bins * tiles, offs.p, st);
        hipLaunchKernelGGL(radix_scatter_kernel, dim3((unsigned)tiles), dim3(RB), 0, st, kin, din, win, kout,
                           dout, wout, R, shift, db, offs.p, tiles);
        SUBWAY_CHECK(hipGetLastError());
        std::swap(kin, kout);
        std::swap(din, dout);
        std::swap(win, wout);
        shift += db;
    }
    counts.reset();
    offs.reset();

    // the sorted records are in the A arrays after an even number of passes
    const bool inA = (passes % 2) == 0;
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
