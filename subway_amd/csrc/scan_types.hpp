// scan_types.hpp -- host-visible control structures of the single-pass
// scan kernels (shared by host code and device code).
#ifndef SUBWAY_SCAN_TYPES_HPP
#define SUBWAY_SCAN_TYPES_HPP

#include <cstddef>
#include <cstdint>

namespace subway {
namespace dev {

typedef uint32_t u32;
typedef unsigned long long u64;

// Per-launch bookkeeping of a single-pass scan kernel.  The last workgroup
// to finish resets the tile and done counters and bumps the epoch, so the
// kernel can be relaunched (or replayed from a graph) with no host memset.
struct ScanCtl {
    u32 tile_counter;
    u32 done_counter;
    u32 epoch;
    u32 stop;       // frontier kernels: previous launch emitted nothing
    u64 count;      // number of selected items of the last launch
    u64 total;      // sum of the selected items' weights
    u32 error;      // nonzero: a bounded device wait gave up (WAIT_* code);
    u32 pad0;       // count, total, error form the [count, total, error] totals block
    u32 adv_counter;  // chunk counter of the consuming advance kernel
    u32 adv_done;
    u64 updates;    // successful relaxations (statistics)
    u64 rounds;     // launches that selected at least one item
    u64 entry_sum;  // selected items over all launches (statistics)
    u64 work_sum;   // selected weight over all launches (statistics)
    u64 item_sum;   // items scanned over all launches (statistics)
    // generator only: out-degree histogram of the selected items (< 4, 4..31,
    // 32..255, >= 256), accumulated until the scheduler's policy kernel reads
    // and clears it
    u64 hist[4];
    // batched frontiers: (edge, active source) pairs and active sources of
    // the selected items, over all launches (statistics)
    u64 lane_work;
    u64 lane_entries;
    // the part of entry_sum, work_sum and updates that stand-alone
    // (non-persistent) frontier and advance launches contributed
    u64 pair_entries;
    u64 pair_work;
    u64 pair_updates;
    // bottom-up BFS steps: vertices found and their out-edges (the next
    // frontier), cleared by the host before each step
    u64 pull_found;
    u64 pull_edges;
    // the part of entry_sum, work_sum and updates that persistent rounds
    // launches over units in the block layout contributed
    u64 blk_entries;
    u64 blk_work;
    u64 blk_updates;
};
static_assert(offsetof(ScanCtl, lane_entries) == offsetof(ScanCtl, lane_work) + 8,
              "lane_work, lane_entries form one two-word block");
static_assert(offsetof(ScanCtl, total) == offsetof(ScanCtl, count) + 8 &&
                  offsetof(ScanCtl, error) == offsetof(ScanCtl, count) + 16,
              "count, total and error form one [count, total, error] block of u64 words");


// Control block of the persistent rounds kernel: a two-level grid barrier
// (8 group counters on separate cache lines, then one top counter), and the
// tile / chunk claim counters of the two phases, double-buffered by round
// parity so they can be reset one round ahead.
struct PersistCtl {
    u32 bar_sub[8 * 16];
    u32 bar_top[16];
    u32 bar_gen[16];
    u32 tile_ctr[2];
    u32 chunk_ctr[2];
    u32 rounds_done;   // rounds that selected work in the last launch
    u32 converged;     // 1: the last launch ended on a round that selected nothing
    u32 error;         // nonzero: a bounded device wait gave up (WAIT_* code)
    // membership of the running launch: arrivals (bit 31 set once the first
    // workgroup has closed membership), the member count it published, and
    // workgroups that have left (the last one clears all three)
    u32 join;
    u32 members;
    u32 left;
    u32 min_members;   // fewest members of any launch since the host last reset it
    // first barrier wait that gave up: round << 1 | phase (0 frontier, 1
    // advance), members, arrivals counted at its group and top counters,
    // grid size
    u32 diag[5];
};

// Every device-side wait (lookback, grid barrier) polls at most a bounded
// number of times; on expiry it records one of these codes in the error word
// of its control block and its kernel runs to its end without further waits.
constexpr u32 WAIT_LOOKBACK = 1;
constexpr u32 WAIT_BARRIER = 2;
constexpr u32 WAIT_DATA = 3;       // a frontier consumer found inconsistent entries / chunk map
constexpr u32 SPIN_LIMIT_DEFAULT = 1u << 22;   // polls (each a device-scope load + s_sleep)

struct TileDesc {
    u64 *flag;     // per tile: epoch | state | selected count
    u64 *inc;      // per tile: epoch | state | weight sum
};

}  // namespace dev
}  // namespace subway

#endif  // SUBWAY_SCAN_TYPES_HPP
