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

// --------------------------------------------- epoch-tagged tile descriptor
// word = epoch(30) | state(2) | count(32).  A descriptor whose epoch differs
// from the running launch's epoch reads as "not yet published".
constexpr u32 ST_NONE = 0, ST_AGG = 1, ST_INC = 2;
constexpr u32 EPOCH_MASK = 0x3FFFFFFFu;

__device__ __forceinline__ u64 pack_desc(u32 epoch, u32 state, u32 count)
{
    return ((u64)(epoch & EPOCH_MASK) << 34) | ((u64)state << 32) | (u64)count;
}
__device__ __forceinline__ u32 desc_epoch(u64 w) { return (u32)(w >> 34); }
__device__ __forceinline__ u32 desc_state(u64 w) { return (u32)(w >> 32) & 3u; }
__device__ __forceinline__ u32 desc_count(u64 w) { return (u32)w; }

// Decoupled lookback, executed by one full wave of the workgroup.  Publishes
// the tile's aggregate, walks back 64 predecessors at a time, and publishes
// the inclusive prefix.  Returns the exclusive (count, weight) prefix on every
// lane of the calling wave.
__device__ __forceinline__ void lookback(const TileDesc &td, u32 tile, u32 epoch,
                                         u32 tile_cnt, u64 tile_sum,
                                         u32 &excl_cnt, u64 &excl_sum)
{
    const int lane = lane_id();
    if (tile == 0) {
        if (lane == 0) {
            st_relaxed(&td.inc[0], tile_sum);
            st_release(&td.flag[0], pack_desc(epoch, ST_INC, tile_cnt));
        }
        excl_cnt = 0;
        excl_sum = 0;
        return;
    }
    if (lane == 0) {
        st_relaxed(&td.agg[tile], tile_sum);
        st_release(&td.flag[tile], pack_desc(epoch, ST_AGG, tile_cnt));
    }
    u32 run_cnt = 0;
    u64 run_sum = 0;
    int64_t base = (int64_t)tile - 1;
    const u32 ep = epoch & EPOCH_MASK;
    for (;;) {
        int64_t idx = base - lane;
        u64 w = 0;
        u32 st = ST_INC;  // lanes before tile 0 read as an empty inclusive prefix
        if (idx >= 0) {
            for (;;) {
                w = ld_acquire(&td.flag[idx]);
                if (desc_epoch(w) == ep && desc_state(w) != ST_NONE) break;
                __builtin_amdgcn_s_sleep(1);
            }
            st = desc_state(w);
        }
        unsigned long long incm = __ballot(st == ST_INC);
        // incm is never zero once idx < 0 appears, so the walk terminates
        int first = incm ? (__ffsll((long long)incm) - 1) : WAVE;
        u32 c = 0;
        u64 s = 0;
        if (lane <= first && idx >= 0) {
            c = desc_count(w);
            s = (lane == first) ? ld_relaxed(&td.inc[idx]) : ld_relaxed(&td.agg[idx]);
        }
        run_cnt += wave_reduce(c);
        run_sum += wave_reduce(s);
        if (incm) break;
        base -= WAVE;
    }
    if (lane == 0) {
        st_relaxed(&td.inc[tile], run_sum + tile_sum);
        st_release(&td.flag[tile], pack_desc(epoch, ST_INC, run_cnt + tile_cnt));
    }
    excl_cnt = run_cnt;
    excl_sum = run_sum;
}

}  // namespace dev
}  // namespace subway

#endif  // SUBWAY_DEVICE_HPP
