// radix_sort.hpp -- device sort and scan primitives of the GPU graph builds
// (R-MAT generation, vertex renumbering): a stable LSD radix sort of
// (key, payload, optional byte) records by key and an exclusive scan of u32
// counts into u64 offsets.
//
// Per digit pass of the sort: tile histograms stored digit-major, one
// exclusive scan over them (so every tile knows where each digit's records
// of that tile start), and a stable scatter whose in-tile ranks come from
// wave64 ballots.  Equal keys keep their input order.
#ifndef SUBWAY_RADIX_SORT_HPP
#define SUBWAY_RADIX_SORT_HPP

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <utility>

#include "device.hpp"
#include "subway/globals.hpp"
#include "subway/gpu_error_check.hpp"

namespace subway {
namespace {

using namespace dev;

constexpr int RB = 256;                 // threads per workgroup of the sort kernels
constexpr int RIPT = 32;                // records per thread and tile
constexpr int RTILE = RB * RIPT;        // records per tile
constexpr int NWAVE = RB / WAVE;
constexpr int SCAN_IPT = 16;
constexpr int SCAN_TILE = RB * SCAN_IPT;

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
            throw Error(ERR_NOMEM, "device allocation of " + std::to_string(n * sizeof(T)) + " bytes failed (" +
                                       hipGetErrorString(e) + ")");
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

// Stable sort of R records (key, val, w) by the low keyBits bits of key,
// digits of at most 8 bits spread evenly over the passes.  The B arrays
// are scratch of the same sizes; on return the sorted records are in
// whichever pair the pointers name (swapped in place).  w may be null.
inline void sort_records(u32 *&key, u32 *&val, uint8_t *&w, u32 *&keyB, u32 *&valB, uint8_t *&wB, u64 R,
                         int keyBits, hipStream_t st)
{
    const int passes = (keyBits + 7) / 8;
    const u64 tiles = (R + RTILE - 1) / RTILE;
    DevArr<u32> counts(tiles * 256);
    DevArr<u64> offs(tiles * 256);
    int shift = 0;
    for (int ps = 0; ps < passes; ps++) {
        const int db = (keyBits - shift + (passes - ps) - 1) / (passes - ps);
        const int bins = 1 << db;
        hipLaunchKernelGGL(radix_hist_kernel, dim3((unsigned)tiles), dim3(RB), 0, st, key, R, shift, bins,
                           counts.p, tiles);
        SUBWAY_CHECK(hipGetLastError());
        scan_counts(counts.p, (u64)bins * tiles, offs.p, st);
        hipLaunchKernelGGL(radix_scatter_kernel, dim3((unsigned)tiles), dim3(RB), 0, st, key, val, w, keyB, valB,
                           wB, R, shift, db, offs.p, tiles);
        SUBWAY_CHECK(hipGetLastError());
        std::swap(key, keyB);
        std::swap(val, valB);
        std::swap(w, wB);
        shift += db;
    }
    SUBWAY_CHECK(hipStreamSynchronize(st));
}

}  // namespace
}  // namespace subway

#endif  // SUBWAY_RADIX_SORT_HPP
