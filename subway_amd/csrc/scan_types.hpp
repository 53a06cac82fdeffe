// scan_types.hpp -- host-visible control structures of the single-pass
// scan kernels (shared by host code and device code).
#ifndef SUBWAY_SCAN_TYPES_HPP
#define SUBWAY_SCAN_TYPES_HPP

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
    u32 adv_counter;  // chunk counter of the consuming advance kernel
    u32 adv_done;
    u64 updates;    // successful relaxations (statistics)
    u64 rounds;     // launches that selected at least one item
    u64 entry_sum;  // selected items over all launches (statistics)
    u64 work_sum;   // selected weight over all launches (statistics)
};

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
};

struct TileDesc {
    u64 *flag;     // packed descriptors
    u64 *agg;      // weight aggregate of the tile
    u64 *inc;      // inclusive weight prefix through the tile
};

}  // namespace dev
}  // namespace subway

#endif  // SUBWAY_SCAN_TYPES_HPP
