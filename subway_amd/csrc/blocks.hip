// blocks.hip -- builds a graph's block-compressed rows (layout: blocks.hpp) on
// the GPU.
//
// Steps, all on the device:
//   1. the edges as the host holds them (split arrays, the 8-byte or 4-byte
//      list) are uploaded as destinations + weights, and the weight span is
//      reduced to (wbase, wbits);
//   2. every row is sorted by destination: chunks of whole rows of at most
//      2^29 edges, each a stable radix sort of (destination, position)
//      records by destination and then of (row, position) records by row,
//      which leaves the positions of every row in destination order; the
//      edges are gathered through them;
//   3. each host layout the graph holds (split arrays, 8-byte list, packed
//      words, 4-byte list) is rewritten from the sorted rows;
//   4. blocks per row (ceil(degree / 32)) and their exclusive scan; per
//      block its gap width and payload words, their exclusive scan; the
//      row offsets rowPtr[v] = 2 * (blocks before v) + (payload words before
//      v);
//   5. one thread per block writes its two descriptor words and its payload
//      words (blocks own whole words, so no two threads write one word);
//   6. the words go to pinned host memory, rowPtr to ordinary memory.
// A vertex's out-edges keep their multiset, so every application's result
// is unchanged by the sort.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "blocks.hpp"
#include "device.hpp"
#include "host_util.hpp"
#include "radix_sort.hpp"
#include "subway/globals.hpp"
#include "subway/gpu_error_check.hpp"
#include "subway/graph.hpp"

namespace subway {

using namespace dev;

namespace {

constexpr int BB = 256;                 // threads per workgroup
constexpr int BTILE = BB * 8;           // blocks per workgroup of the block kernels
constexpr u64 SORT_CHUNK = 1ull << 29;  // edges per row chunk of the sort

inline unsigned grid_of(u64 items)
{
    return (unsigned)std::max<u64>(1, std::min<u64>((items + BB - 1) / BB, 1u << 16));
}

// Copies between the device and host edge arrays.  Pinned arrays are copied
// directly; arrays in ordinary memory go through a pinned bounce buffer in
// chunks, so that the runtime never has to lock existing host pages itself
// (registering existing memory is refused on some hosts).
class HostCopy {
public:
    explicit HostCopy(hipStream_t st) : st_(st) {}
    ~HostCopy()
    {
        if (buf_) (void)hipHostFree(buf_);
    }
    void h2d(void *dev, const void *host, u64 bytes, bool pinned)
    {
        if (!bytes) return;
        if (pinned) {
            SUBWAY_CHECK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, st_));
            return;
        }
        bounce();
        for (u64 a = 0; a < bytes; a += kBounce) {
            const u64 n = std::min<u64>(kBounce, bytes - a);
            parallel_for(n, [&](u64 b, u64 e, int) { std::memcpy(buf_ + b, (const char *)host + a + b, e - b); });
            SUBWAY_CHECK(hipMemcpyAsync((char *)dev + a, buf_, n, hipMemcpyHostToDevice, st_));
            SUBWAY_CHECK(hipStreamSynchronize(st_));
        }
    }
    void d2h(void *host, const void *dev, u64 bytes, bool pinned)
    {
        if (!bytes) return;
        if (pinned) {
            SUBWAY_CHECK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st_));
            return;
        }
        bounce();
        for (u64 a = 0; a < bytes; a += kBounce) {
            const u64 n = std::min<u64>(kBounce, bytes - a);
            SUBWAY_CHECK(hipMemcpyAsync(buf_, (const char *)dev + a, n, hipMemcpyDeviceToHost, st_));
            SUBWAY_CHECK(hipStreamSynchronize(st_));
            parallel_for(n, [&](u64 b, u64 e, int) { std::memcpy((char *)host + a + b, buf_ + b, e - b); });
        }
    }

private:
    static constexpr u64 kBounce = 64ull << 20;
    hipStream_t st_;
    char *buf_ = nullptr;
    void bounce()
    {
        if (!buf_) SUBWAY_CHECK(hipHostMalloc((void **)&buf_, kBounce, 0));
    }
};

// last r in [lo, hi] with a[r] <= q (a[lo] <= q holds)
__device__ __forceinline__ u64 last_le(const u64 *a, u64 lo, u64 hi, u64 q)
{
    while (lo < hi) {
        const u64 mid = lo + (hi - lo + 1) / 2;
        if (a[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// destinations (and u32 weights) out of an uploaded 8-byte edge list
__global__ __launch_bounds__(BB) void unzip_kernel(const uint2 *e, u64 m, u32 *dst, u32 *w)
{
    const u64 step = (u64)gridDim.x * BB;
    for (u64 i = (u64)blockIdx.x * BB + threadIdx.x; i < m; i += step) {
        const uint2 x = e[i];
        dst[i] = x.x;
        w[i] = x.y;
    }
}

// smallest and largest weight (out[0] = min, out[1] = max; preset to ~0, 0)
template <class W>
__global__ __launch_bounds__(BB) void weight_span_kernel(const W *w, u64 m, u32 *out)
{
    u32 lo = 0xFFFFFFFFu, hi = 0;
    const u64 step = (u64)gridDim.x * BB;
    for (u64 i = (u64)blockIdx.x * BB + threadIdx.x; i < m; i += step) {
        const u32 x = (u32)w[i];
        lo = x < lo ? x : lo;
        hi = x > hi ? x : hi;
    }
    lo = wave_reduce_min(lo);
    hi = wave_reduce_max(hi);
    if (lane_id() == 0) {
        atomicMin(&out[0], lo);
        atomicMax(&out[1], hi);
    }
}

// sort records of a chunk: key = destination, value = position in the chunk
__global__ __launch_bounds__(BB) void dst_keys_kernel(const u32 *dst, u64 e0, u64 cnt, u32 *key, u32 *pos)
{
    const u64 step = (u64)gridDim.x * BB;
    for (u64 i = (u64)blockIdx.x * BB + threadIdx.x; i < cnt; i += step) {
        key[i] = dst[e0 + i];
        pos[i] = (u32)i;
    }
}

// key[i] = row of chunk edge pos[i], relative to the chunk's first row r0
__global__ __launch_bounds__(BB) void row_keys_kernel(const u64 *ptr, u64 r0, u64 r1, u64 e0, const u32 *pos,
                                                      u64 cnt, u32 *key)
{
    const u64 step = (u64)gridDim.x * BB;
    for (u64 i = (u64)blockIdx.x * BB + threadIdx.x; i < cnt; i += step) {
        const u64 e = e0 + pos[i];
        key[i] = (u32)(last_le(ptr, r0, r1 - 1, e) - r0);
    }
}

template <class W>
__global__ __launch_bounds__(BB) void permute_kernel(const u32 *dst, const W *w, u64 e0, const u32 *pos, u64 cnt,
                                                     u32 *dOut, W *wOut)
{
    const u64 step = (u64)gridDim.x * BB;
    for (u64 i = (u64)blockIdx.x * BB + threadIdx.x; i < cnt; i += step) {
        const u64 s = e0 + pos[i];
        dOut[i] = dst[s];
        if (w) wOut[i] = w[s];
    }
}

// host layouts rebuilt from the sorted rows
template <class W>
__global__ __launch_bounds__(BB) void zip_kernel(const u32 *dst, const W *w, u64 e0, u64 cnt, uint2 *out)
{
    const u64 step = (u64)gridDim.x * BB;
    for (u64 i = (u64)blockIdx.x * BB + threadIdx.x; i < cnt; i += step)
        out[i] = make_uint2(dst[e0 + i], (u32)w[e0 + i]);
}
template <class W>
__global__ __launch_bounds__(BB) void pack_kernel(const u32 *dst, const W *w, u64 e0, u64 cnt, u32 shift, u32 base,
                                                  u32 *out)
{
    const u64 step = (u64)gridDim.x * BB;
    for (u64 i = (u64)blockIdx.x * BB + threadIdx.x; i < cnt; i += step)
        out[i] = dst[e0 + i] | (((u32)w[e0 + i] - base) << shift);
}
template <class W>
__global__ __launch_bounds__(BB) void narrow_kernel(const W *w, u64 e0, u64 cnt, uint8_t *out)
{
    const u64 step = (u64)gridDim.x * BB;
    for (u64 i = (u64)blockIdx.x * BB + threadIdx.x; i < cnt; i += step) out[i] = (uint8_t)w[e0 + i];
}

// blocks per row; nb[n] = 0 so that the scan's last entry is the total
__global__ __launch_bounds__(BB) void row_blocks_kernel(const u64 *ptr, u64 n, u32 *nb)
{
    const u64 step = (u64)gridDim.x * BB;
    for (u64 v = (u64)blockIdx.x * BB + threadIdx.x; v <= n; v += step)
        nb[v] = v < n ? (u32)((ptr[v + 1] - ptr[v] + BLK_EDGES - 1) >> BLK_SHIFT) : 0u;
}

// The row of each block of a tile [t * BTILE, (t + 1) * BTILE): thread 0
// bounds the tile's rows, every block then searches that range only.
struct BlockSpot {
    u64 v, k, e;
    u32 cnt, nb;
};
__device__ __forceinline__ BlockSpot block_spot(const u64 *ptr, const u64 *bidx, u64 b, u64 p0, u64 p1)
{
    BlockSpot s;
    s.v = last_le(bidx, p0, p1, b);
    s.k = b - bidx[s.v];
    const u64 r0 = ptr[s.v], d = ptr[s.v + 1] - r0;
    s.e = r0 + (s.k << BLK_SHIFT);
    const u64 left = d - (s.k << BLK_SHIFT);
    s.cnt = (u32)(left < (u64)BLK_EDGES ? left : (u64)BLK_EDGES);
    s.nb = (u32)((d + BLK_EDGES - 1) >> BLK_SHIFT);
    return s;
}

// bits of the widest gap between consecutive (sorted) destinations of a block
__device__ __forceinline__ u32 block_width(const u32 *d, u32 cnt)
{
    u32 mx = 0;
    for (u32 j = 1; j < cnt; j++) {
        const u32 g = d[j] - d[j - 1];
        mx = g > mx ? g : mx;
    }
    return blk_bits(mx);
}

__global__ __launch_bounds__(BB) void block_words_kernel(const u64 *ptr, const u64 *bidx, u64 n, u64 B,
                                                         const u32 *dst, u32 wbits, u32 *pw)
{
    __shared__ u64 s_p0, s_p1;
    const u64 b0 = (u64)blockIdx.x * BTILE;
    const u64 b1 = b0 + BTILE < B ? b0 + BTILE : B;
    if (threadIdx.x == 0) {
        s_p0 = last_le(bidx, 0, n - 1, b0);
        s_p1 = last_le(bidx, s_p0, n - 1, b1 - 1);
    }
    __syncthreads();
    for (u64 b = b0 + threadIdx.x; b < b1; b += BB) {
        const BlockSpot s = block_spot(ptr, bidx, b, s_p0, s_p1);
        pw[b] = blk_block_words(s.cnt, block_width(dst + s.e, s.cnt), wbits, s.nb == 1);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) pw[B] = 0;
}

// pay: exclusive prefix of the blocks' words (descriptors included), so a
// row starts where its first block does
__global__ __launch_bounds__(BB) void row_ptr_kernel(const u64 *bidx, const u64 *pay, u64 n, u64 *rowPtr)
{
    const u64 step = (u64)gridDim.x * BB;
    for (u64 v = (u64)blockIdx.x * BB + threadIdx.x; v <= n; v += step) rowPtr[v] = pay[bidx[v]];
}

template <class W>
__global__ __launch_bounds__(BB) void encode_kernel(const u64 *ptr, const u64 *bidx, const u64 *pay,
                                                    const u64 *rowPtr, u64 n, u64 B, const u32 *dst, const W *w,
                                                    u32 wbits, u32 wbase, u32 *out, u32 *err)
{
    __shared__ u64 s_p0, s_p1;
    const u64 b0 = (u64)blockIdx.x * BTILE;
    const u64 b1 = b0 + BTILE < B ? b0 + BTILE : B;
    if (threadIdx.x == 0) {
        s_p0 = last_le(bidx, 0, n - 1, b0);
        s_p1 = last_le(bidx, s_p0, n - 1, b1 - 1);
    }
    __syncthreads();
    for (u64 b = b0 + threadIdx.x; b < b1; b += BB) {
        const BlockSpot s = block_spot(ptr, bidx, b, s_p0, s_p1);
        const u32 anchor = dst[s.e];
        const u32 wd = block_width(dst + s.e, s.cnt);
        const u64 row = rowPtr[s.v];
        u64 at, acc = 0;
        u32 used = 0;
        if (s.nb == 1) {
            // one block: the anchor, then the payload led by the width
            out[row] = anchor;
            at = row + 1;
            acc = wd;
            used = BLK_WFIELD;
        } else {
            // words of the row's earlier blocks, less their descriptors
            const u64 woff = pay[b] - pay[bidx[s.v]] - 2 * s.k;
            if (woff + (pay[b + 1] - pay[b] - 2) > (u64)BLK_MAX_ROW_WORDS) {
                atomicOr(err, 1u);
                continue;
            }
            out[row + 2 * s.k] = anchor;
            out[row + 2 * s.k + 1] = ((u32)woff << BLK_WFIELD) | wd;
            at = row + 2ull * s.nb + woff;
        }
        auto put = [&](u32 x, u32 width) {
            if (width == 0) return;
            acc |= (u64)x << used;
            used += width;
            if (used >= 32) {
                out[at++] = (u32)acc;
                acc >>= 32;
                used -= 32;
            }
        };
        for (u32 j = 1; j < s.cnt; j++) put(dst[s.e + j] - dst[s.e + j - 1], wd);
        if (w)
            for (u32 j = 0; j < s.cnt; j++) put((u32)w[s.e + j] - wbase, wbits);
        if (used) out[at] = (u32)acc;
    }
}

template <class W>
void sort_rows(const u64 *hptr, const u64 *dptr, u64 n, u64 maxDeg, u32 *dst, W *w, hipStream_t st)
{
    const u64 m = hptr[n];
    if (m == 0) return;
    const u64 C = std::min<u64>(m, std::max<u64>(SORT_CHUNK, maxDeg));
    if (C > 0xFFFFFFFFull) throw Error(ERR_ARG, "a row of more than 2^32 edges cannot be sorted here");
    DevArr<u32> kA(C), vA(C), kB(C), vB(C), tD(C);
    DevArr<W> tW;
    if (w) tW.alloc(C);
    const int dbits = std::max<int>(1, (int)blk_bits((u32)(n > 0 ? n - 1 : 0)));
    u64 r0 = 0;
    while (r0 < n) {
        // whole rows [r0, r1) with at most C edges
        u64 r1 = r0 + 1;
        {
            u64 lo = r0 + 1, hi = n;
            while (lo < hi) {
                const u64 mid = lo + (hi - lo + 1) / 2;
                if (hptr[mid] - hptr[r0] <= C) lo = mid;
                else hi = mid - 1;
            }
            r1 = lo;
        }
        const u64 e0 = hptr[r0], cnt = hptr[r1] - e0;
        if (cnt > 1) {
            u32 *key = kA.p, *pos = vA.p, *key2 = kB.p, *pos2 = vB.p;
            uint8_t *nw = nullptr, *nw2 = nullptr;
            hipLaunchKernelGGL(dst_keys_kernel, dim3(grid_of(cnt)), dim3(BB), 0, st, dst, e0, cnt, key, pos);
            SUBWAY_CHECK(hipGetLastError());
            sort_records(key, pos, nw, key2, pos2, nw2, cnt, dbits, st);
            hipLaunchKernelGGL(row_keys_kernel, dim3(grid_of(cnt)), dim3(BB), 0, st, dptr, r0, r1, e0, pos, cnt, key);
            SUBWAY_CHECK(hipGetLastError());
            const int rbits = std::max<int>(1, (int)blk_bits((u32)(r1 - r0 - 1)));
            sort_records(key, pos, nw, key2, pos2, nw2, cnt, rbits, st);
            hipLaunchKernelGGL((permute_kernel<W>), dim3(grid_of(cnt)), dim3(BB), 0, st, dst, (const W *)w, e0, pos,
                               cnt, tD.p, tW.p);
            SUBWAY_CHECK(hipGetLastError());
            SUBWAY_CHECK(hipMemcpyAsync(dst + e0, tD.p, cnt * sizeof(u32), hipMemcpyDeviceToDevice, st));
            if (w) SUBWAY_CHECK(hipMemcpyAsync(w + e0, tW.p, cnt * sizeof(W), hipMemcpyDeviceToDevice, st));
        }
        r0 = r1;
    }
    SUBWAY_CHECK(hipStreamSynchronize(st));
}

// host layouts rewritten from the sorted device rows, in chunks through a
// device staging buffer
template <class E, class W>
void write_back(HostCsr<E> &csr, const u32 *dst, const W *w, hipStream_t st)
{
    constexpr bool weighted = sizeof(E) == 8;
    const u64 m = csr.m;
    if (m == 0) return;
    HostCopy hc(st);
    if constexpr (weighted) {
        if (csr.splitDst) {
            hc.d2h(csr.splitDst, dst, m * sizeof(u32), csr.splitPinned);
            if (sizeof(W) == 1) hc.d2h(csr.splitW, w, m, csr.splitPinned);
        }
        const u64 CH = std::min<u64>(m, 1ull << 27);
        DevArr<uint2> stage(CH);
        for (u64 e0 = 0; e0 < m; e0 += CH) {
            const u64 cnt = std::min(CH, m - e0);
            if (csr.splitDst && sizeof(W) != 1) {
                hipLaunchKernelGGL((narrow_kernel<W>), dim3(grid_of(cnt)), dim3(BB), 0, st, w, e0, cnt,
                                   reinterpret_cast<uint8_t *>(stage.p));
                hc.d2h(csr.splitW + e0, stage.p, cnt, csr.splitPinned);
                SUBWAY_CHECK(hipStreamSynchronize(st));
            }
            if (csr.edges) {
                hipLaunchKernelGGL((zip_kernel<W>), dim3(grid_of(cnt)), dim3(BB), 0, st, dst, w, e0, cnt, stage.p);
                hc.d2h(csr.edges + e0, stage.p, cnt * sizeof(E), csr.pinned);
                SUBWAY_CHECK(hipStreamSynchronize(st));
            }
            if (csr.packed) {
                hipLaunchKernelGGL((pack_kernel<W>), dim3(grid_of(cnt)), dim3(BB), 0, st, dst, w, e0, cnt,
                                   csr.packShift, csr.packBase, reinterpret_cast<u32 *>(stage.p));
                SUBWAY_CHECK(hipMemcpyAsync(csr.packed + e0, stage.p, cnt * sizeof(u32), hipMemcpyDeviceToHost, st));
                SUBWAY_CHECK(hipStreamSynchronize(st));
            }
            SUBWAY_CHECK(hipGetLastError());
        }
    } else {
        (void)w;
        hc.d2h(csr.edges, dst, m * sizeof(u32), csr.pinned);
    }
    SUBWAY_CHECK(hipStreamSynchronize(st));
}

template <class E, class W>
void build(HostCsr<E> &csr, DevArr<u32> &dst, DevArr<W> &w, bool hasW, bool unpinSplit, hipStream_t st)
{
    const u64 n = csr.n, m = csr.m;
    HostCopy hc(st);
    DevArr<u64> dptr(n + 1);
    hc.h2d(dptr.p, csr.ptr, (n + 1) * sizeof(u64), false);

    // 1. weight code
    u32 wbits = 0, wbase = 0;
    if (hasW && m) {
        DevArr<u32> span(2);
        const u32 init[2] = {0xFFFFFFFFu, 0u};
        SUBWAY_CHECK(hipMemcpyAsync(span.p, init, sizeof(init), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL((weight_span_kernel<W>), dim3(grid_of(m)), dim3(BB), 0, st, w.p, m, span.p);
        SUBWAY_CHECK(hipGetLastError());
        u32 got[2];
        SUBWAY_CHECK(hipMemcpyAsync(got, span.p, sizeof(got), hipMemcpyDeviceToHost, st));
        SUBWAY_CHECK(hipStreamSynchronize(st));
        wbase = got[0];
        wbits = blk_bits(got[1] - got[0]);
        if (wbase >= BLK_MAX_WBASE) throw Error(ERR_ARG, "the smallest weight is too large for the block layout");
    }

    // 2-3. rows by destination, host layouts rewritten
    sort_rows<W>(csr.ptr, dptr.p, n, (u64)csr.maxDegree, dst.p, hasW ? w.p : nullptr, st);
    write_back<E, W>(csr, dst.p, hasW ? w.p : nullptr, st);
    csr.rowsSorted = true;

    // 4. blocks, payload words, row offsets
    DevArr<u32> nb(n + 1);
    hipLaunchKernelGGL(row_blocks_kernel, dim3(grid_of(n + 1)), dim3(BB), 0, st, dptr.p, n, nb.p);
    SUBWAY_CHECK(hipGetLastError());
    DevArr<u64> bidx(n + 1);
    scan_counts(nb.p, n + 1, bidx.p, st);
    nb.reset();
    // the build's stream is non-blocking: every read-back goes through it
    // and waits for it (a plain hipMemcpy would not wait for its kernels)
    auto readBack = [&](void *host, const void *dev, size_t bytes) {
        SUBWAY_CHECK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st));
        SUBWAY_CHECK(hipStreamSynchronize(st));
    };
    u64 B = 0;
    readBack(&B, bidx.p + n, sizeof(u64));
    DevArr<u64> rowPtr(n + 1);
    u64 T = 0;
    DevArr<u32> words;
    if (B == 0) {
        SUBWAY_CHECK(hipMemsetAsync(rowPtr.p, 0, (n + 1) * sizeof(u64), st));
    } else {
        DevArr<u32> pw(B + 1);
        const unsigned tiles = (unsigned)((B + BTILE - 1) / BTILE);
        hipLaunchKernelGGL(block_words_kernel, dim3(tiles), dim3(BB), 0, st, dptr.p, bidx.p, n, B, dst.p, wbits, pw.p);
        SUBWAY_CHECK(hipGetLastError());
        DevArr<u64> pay(B + 1);
        scan_counts(pw.p, B + 1, pay.p, st);
        pw.reset();
        hipLaunchKernelGGL(row_ptr_kernel, dim3(grid_of(n + 1)), dim3(BB), 0, st, bidx.p, pay.p, n, rowPtr.p);
        SUBWAY_CHECK(hipGetLastError());
        readBack(&T, rowPtr.p + n, sizeof(u64));

        // 5. descriptors and payloads
        words.alloc(T);
        DevArr<u32> err(1);
        SUBWAY_CHECK(hipMemsetAsync(err.p, 0, sizeof(u32), st));
        hipLaunchKernelGGL((encode_kernel<W>), dim3(tiles), dim3(BB), 0, st, dptr.p, bidx.p, pay.p, rowPtr.p, n, B,
                           dst.p, hasW ? (const W *)w.p : (const W *)nullptr, wbits, wbase, words.p, err.p);
        SUBWAY_CHECK(hipGetLastError());
        u32 e = 0;
        readBack(&e, err.p, sizeof(u32));
        if (e) throw Error(ERR_ARG, "a row is too large for the block layout (2^26 payload words per row)");
    }

    // 6. to the host; when the process cannot pin the words beside pinned
    // split arrays (a host pins about 70 GB per process), the split arrays
    // move to ordinary memory first: the blocks then carry the whole ranges
    // and compacted units are gathered on the host
    (void)unpinSplit;
    u32 *hw = nullptr;
    if (hipHostMalloc((void **)&hw, (size_t)(T ? T : 1) * sizeof(u32), hipHostMallocMapped | hipHostMallocPortable) != hipSuccess) {
        (void)hipGetLastError();
        hw = nullptr;
        if (csr.splitDst && csr.splitPinned) {
            csr.demoteSplit();
            // the runtime releases freed pinned ranges once the device is idle;
            // an allocation made before that gave host pointers that copies
            // then refused ("invalid argument")
            SUBWAY_CHECK(hipDeviceSynchronize());
            if (hipHostMalloc((void **)&hw, (size_t)(T ? T : 1) * sizeof(u32), hipHostMallocMapped | hipHostMallocPortable) != hipSuccess) {
                (void)hipGetLastError();
                hw = nullptr;
            }
        }
        if (!hw)
            throw Error(ERR_NOMEM, "pinned host allocation of " + std::to_string(T * 4) + " bytes of blocks failed");
    }
    ull *hp = nullptr;
    try {
        hp = new ull[n + 1];
    } catch (const std::bad_alloc &) {
        (void)hipHostFree(hw);
        throw Error(ERR_NOMEM, "host allocation of the block row offsets failed");
    }
    try {
        hc.d2h(hw, words.p, T * sizeof(u32), true);
        hc.d2h(hp, rowPtr.p, (n + 1) * sizeof(u64), false);
        SUBWAY_CHECK(hipStreamSynchronize(st));
    } catch (...) {
        (void)hipHostFree(hw);
        delete[] hp;
        throw;
    }
    csr.blk = hw;
    csr.blkPtr = hp;
    csr.blkWords = T;
    csr.blkCode = blk_code(wbits, wbase);
}

// Split arrays in ordinary memory: a pinned copy of the edges of the id
// suffix [v, n) for the smallest v at which the suffix's rows take at least
// their split bytes in the block layout (there whole ranges move in the
// split format, through this copy), at most a quarter of the edges and 2^31
// of them; none when pinned memory runs short.
template <class E>
void pin_split_tail(HostCsr<E> &csr)
{
    const u64 n = csr.n, m = csr.m;
    if (!csr.splitDst || csr.splitPinned || !csr.blkPtr || m == 0) return;
    const ull *np = csr.ptr, *bp = csr.blkPtr;
    const u64 maxTail = std::min<u64>(m / 4, 1ull << 31);
    u64 start = n;
    for (u64 v = n; v-- > 0;) {
        const u64 edges = np[n] - np[v];
        if (edges > maxTail) break;
        if ((bp[n] - bp[v]) * 4 >= edges * 5) start = v;   // 4 bytes a word, 5 an edge
    }
    const u64 t0 = np[start], cnt = m - t0;
    if (cnt == 0) return;
    uint32_t *d = nullptr;
    uint8_t *w = nullptr;
    if (hipHostMalloc((void **)&d, cnt * sizeof(uint32_t), hipHostMallocPortable) != hipSuccess ||
        hipHostMalloc((void **)&w, cnt, hipHostMallocPortable) != hipSuccess) {
        (void)hipGetLastError();
        if (d) (void)hipHostFree(d);
        return;
    }
    parallel_for(cnt, [&](u64 b, u64 e, int) {
        std::memcpy(d + b, csr.splitDst + t0 + b, (e - b) * sizeof(uint32_t));
        std::memcpy(w + b, csr.splitW + t0 + b, e - b);
    });
    csr.tailDst = d;
    csr.tailW = w;
    csr.tailEdge = t0;
}

}  // namespace

template <class E>
void build_blocks(HostCsr<E> &csr, bool unpinSplit)
{
    constexpr bool weighted = sizeof(E) == 8;
    if (csr.blk) throw Error(ERR_ARG, "the graph already holds its block layout");
    const u64 n = csr.n, m = csr.m;
    if (n == 0) throw Error(ERR_ARG, "the block layout needs at least one vertex");
    if (m > 0 && !csr.edges && !csr.splitDst) throw Error(ERR_ARG, "the graph holds no edge list or split arrays");
    hipStream_t st = nullptr;
    SUBWAY_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    // split arrays asked to leave pinned memory leave it first: the uploads and
    // write-backs below then go through the bounce buffer, and the device has
    // long been idle when the block words are pinned
    if (unpinSplit && csr.splitDst && csr.splitPinned) {
        csr.demoteSplit();
        SUBWAY_CHECK(hipDeviceSynchronize());
    }
    DevArr<u32> dst(m);
    HostCopy hc(st);
    if constexpr (weighted) {
        if (csr.splitDst) {
            DevArr<uint8_t> w(m);
            hc.h2d(dst.p, csr.splitDst, m * sizeof(u32), csr.splitPinned);
            hc.h2d(w.p, csr.splitW, m, csr.splitPinned);
            SUBWAY_CHECK(hipStreamSynchronize(st));
            build<E, uint8_t>(csr, dst, w, true, unpinSplit, st);
        } else {
            DevArr<u32> w(m);
            if (m) {
                DevArr<uint2> list(m);
                hc.h2d(list.p, csr.edges, m * sizeof(E), csr.pinned);
                hipLaunchKernelGGL(unzip_kernel, dim3(grid_of(m)), dim3(BB), 0, st, list.p, m, dst.p, w.p);
                SUBWAY_CHECK(hipGetLastError());
                SUBWAY_CHECK(hipStreamSynchronize(st));
            }
            build<E, u32>(csr, dst, w, true, unpinSplit, st);
        }
    } else {
        hc.h2d(dst.p, csr.edges, m * sizeof(u32), csr.pinned);
        SUBWAY_CHECK(hipStreamSynchronize(st));
        DevArr<uint8_t> none;
        build<E, uint8_t>(csr, dst, none, false, unpinSplit, st);
    }
    pin_split_tail(csr);
}

template void build_blocks<OutEdge>(HostCsr<OutEdge> &, bool);
template void build_blocks<OutEdgeWeighted>(HostCsr<OutEdgeWeighted> &, bool);

}  // namespace subway
