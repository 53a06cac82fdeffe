// device.hpp -- wave64 / workgroup primitives for gfx950 used by every
// kernel of the pipeline: wave inclusive scans, a workgroup scan of
// (count, weight) pairs, and a single-pass decoupled-lookback tile prefix
// with epoch-tagged tile descriptors (no per-launch clearing).
#ifndef SUBWAY_DEVICE_HPP
#define SUBWAY_DEVICE_HPP

#include <hip/hip_runtime.h>
#include <cstdint>

#include "scan_types.hpp"

namespace subway {
namespace dev {

constexpr int WAVE = 64;

typedef uint32_t u32;
typedef unsigned long long u64;

// ---------------------------------------------------------------- atomics
__device__ __forceinline__ u64 ld_acquire(const u64 *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 ld_relaxed(const u64 *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 ld_relaxed(const u32 *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release(u64 *p, u64 v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(u64 *p, u64 v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(u32 *p, u32 v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 fetch_add_acq_rel(u32 *p, u32 v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 fetch_add_relaxed(u32 *p, u32 v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Memory policy of data exchanged between workgroups inside one launch.
// COH: agent-scope relaxed atomic loads and stores, performed at the device's
// coherence point (the XCDs' L2 caches are not coherent with each other), so
// a grid barrier needs only completion waits, no L2 write-back or
// invalidate.  Without COH: plain accesses, ordered by kernel boundaries.
template <bool COH, class T>
__device__ __forceinline__ T mld(const T *p)
{
    if constexpr (COH) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *p;
}
template <bool COH, class T>
__device__ __forceinline__ void mst(T *p, T v)
{
    if constexpr (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

// ------------------------------------------------------------- wave scans
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ u32 wave_incl_scan(u32 x)
{
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        u32 y = __shfl_up(x, o, WAVE);
        if (lane >= o) x += y;
    }
    return x;
}

__device__ __forceinline__ u64 wave_incl_scan(u64 x)
{
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        u64 y = __shfl_up((unsigned long long)x, o, WAVE);
        if (lane >= o) x += y;
    }
    return x;
}

__device__ __forceinline__ u64 wave_reduce(u64 x)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor((unsigned long long)x, o, WAVE);
    return x;
}

__device__ __forceinline__ u32 wave_reduce(u32 x)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, WAVE);
    return x;
}

// Inclusive scan of x over the lanes [segStart, lane] of a wave64
// (segStart = first lane of the calling lane's segment, 0 for a plain scan),
// with DPP moves: four row shifts within 16-lane rows, then the last lane of
// row 0 / 2 broadcast into row 1 / 3, then lane 31 into rows 2 and 3.  A lane
// adds a moved value only when its source lane lies in its own segment.
__device__ __forceinline__ u32 wave_seg_incl_scan(u32 x, int lane, int segStart)
{
    const int r = lane & 15;
    u32 y;
    y = (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false);   // row_shr:1
    if (r >= 1 && lane - 1 >= segStart) x += y;
    y = (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false);   // row_shr:2
    if (r >= 2 && lane - 2 >= segStart) x += y;
    y = (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false);   // row_shr:4
    if (r >= 4 && lane - 4 >= segStart) x += y;
    y = (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false);   // row_shr:8
    if (r >= 8 && lane - 8 >= segStart) x += y;
    y = (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);   // row_bcast:15 into rows 1, 3
    if (((lane >> 4) & 1) && (lane & ~15) - 1 >= segStart) x += y;
    y = (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);   // row_bcast:31 into rows 2, 3
    if (lane >= 32 && 31 >= segStart) x += y;
    return x;
}

__device__ __forceinline__ u32 wave_reduce_min(u32 x)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const u32 y = __shfl_xor(x, o, WAVE);
        x = y < x ? y : x;
    }
    return x;
}

__device__ __forceinline__ u32 wave_reduce_max(u32 x)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const u32 y = __shfl_xor(x, o, WAVE);
        x = y > x ? y : x;
    }
    return x;
}

// --------------------------------------------- epoch-tagged tile descriptor
// A tile publishes two self-contained 64-bit words, each epoch(22) |
// state(2) | value(40): one carries the selected count, the other the weight
// sum.  A reader accepts a tile only when both words carry the running
// launch's epoch and the same state, so no fence is needed between them
// (every access is a relaxed agent-scope atomic); a word from an older
// launch reads as "not yet published".
constexpr u32 ST_NONE = 0, ST_AGG = 1, ST_INC = 2;
constexpr u32 EPOCH_BITS = 22;
constexpr u32 EPOCH_MASK = (1u << EPOCH_BITS) - 1u;
constexpr u64 VALUE_MASK = (1ull << 40) - 1ull;

__device__ __forceinline__ u64 pack_desc(u32 epoch, u32 state, u64 value)
{
    return ((u64)(epoch & EPOCH_MASK) << 42) | ((u64)state << 40) | (value & VALUE_MASK);
}
__device__ __forceinline__ u32 desc_epoch(u64 w) { return (u32)(w >> 42); }
__device__ __forceinline__ u32 desc_state(u64 w) { return (u32)(w >> 40) & 3u; }
__device__ __forceinline__ u64 desc_value(u64 w) { return w & VALUE_MASK; }

// Error word and poll limit of a bounded wait.
struct Watch {
    u32 *err;
    u32 limit;
};

// Decoupled lookback, executed by one full wave of the workgroup.  Publishes
// the tile's aggregate, walks back 64 predecessors at a time, and publishes
// the inclusive prefix.  Returns the exclusive (count, weight) prefix on every
// lane of the calling wave.  td.flag holds the count words, td.inc the sum
// words.  Returns false (prefix unusable, WAIT_LOOKBACK recorded) when a
// predecessor stays unpublished for more than wt.limit polls or another wait
// of the launch has already given up.
__device__ __forceinline__ bool lookback(const TileDesc &td, u32 tile, u32 epoch, u32 tile_cnt, u64 tile_sum,
                                         u32 &excl_cnt, u64 &excl_sum, const Watch &wt)
{
    const int lane = lane_id();
    excl_cnt = 0;
    excl_sum = 0;
    if (tile == 0) {
        if (lane == 0) {
            st_relaxed(&td.flag[0], pack_desc(epoch, ST_INC, tile_cnt));
            st_relaxed(&td.inc[0], pack_desc(epoch, ST_INC, tile_sum));
        }
        return true;
    }
    if (lane == 0) {
        st_relaxed(&td.flag[tile], pack_desc(epoch, ST_AGG, tile_cnt));
        st_relaxed(&td.inc[tile], pack_desc(epoch, ST_AGG, tile_sum));
    }
    u64 run_cnt = 0;
    u64 run_sum = 0;
    int64_t base = (int64_t)tile - 1;
    const u32 ep = epoch & EPOCH_MASK;
    for (;;) {
        int64_t idx = base - lane;
        u64 wc = 0, ws = 0;
        u32 st = ST_INC;  // lanes before tile 0 read as an empty inclusive prefix
        bool gaveUp = false;
        if (idx >= 0) {
            u32 polls = 0;
            for (;;) {
                wc = ld_relaxed(&td.flag[idx]);
                ws = ld_relaxed(&td.inc[idx]);
                const u32 sc = desc_state(wc);
                if (desc_epoch(wc) == ep && desc_epoch(ws) == ep && (sc == ST_AGG || sc == ST_INC) &&
                    sc == desc_state(ws))
                    break;
                ++polls;
                if (polls > wt.limit || ((polls & 63u) == 0 && ld_relaxed(wt.err) != 0)) {
                    gaveUp = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            st = desc_state(wc);
        }
        if (__ballot(gaveUp)) {
            if (lane == 0) (void)atomicCAS(wt.err, 0u, WAIT_LOOKBACK);
            return false;
        }
        unsigned long long incm = __ballot(st == ST_INC);
        // incm is never zero once idx < 0 appears, so the walk terminates
        int first = incm ? (__ffsll((long long)incm) - 1) : WAVE;
        u64 c = 0, s = 0;
        if (lane <= first && idx >= 0) {
            c = desc_value(wc);
            s = desc_value(ws);
        }
        run_cnt += wave_reduce(c);
        run_sum += wave_reduce(s);
        if (incm) break;
        base -= WAVE;
    }
    if (lane == 0) {
        st_relaxed(&td.flag[tile], pack_desc(epoch, ST_INC, run_cnt + tile_cnt));
        st_relaxed(&td.inc[tile], pack_desc(epoch, ST_INC, run_sum + tile_sum));
    }
    excl_cnt = (u32)run_cnt;
    excl_sum = run_sum;
    return true;
}

}  // namespace dev
}  // namespace subway

#endif  // SUBWAY_DEVICE_HPP
