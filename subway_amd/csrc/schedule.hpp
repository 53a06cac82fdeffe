// schedule.hpp -- host-side choice of the learned schedule's in-place visits:
// which device buffers still hold the whole edge slice of an id range that
// this global iteration should visit where it lies, with no transfer.
#ifndef SUBWAY_SCHEDULE_HPP
#define SUBWAY_SCHEDULE_HPP

#include <algorithm>
#include <cstddef>
#include <cstdint>

namespace subway {

struct ResidentChoice {
    bool visit[2] = {false, false};   // buffer b's slice is visited in place
    uint32_t overlaps = 0;            // slices skipped because an overlapping slice was already chosen
};

// bounds[0..P]: static id-range bounds (range i is [bounds[i], bounds[i+1]));
// activeEdges(i): this iteration's active edges of range i; sliceLo/Hi[b]:
// the id range whose whole edge slice buffer b holds (hi <= lo: none).  A
// slice is visited when it is made of whole ranges, shares no range with a
// slice chosen before it (both buffers can hold overlapping slices, moved
// there in different iterations: the overlap is visited once), and its
// active edges come to minBytes or more.  done[0..P) marks the ranges of
// the chosen slices.
template <class ActiveEdges>
ResidentChoice choose_resident_visits(const uint32_t *bounds, size_t P, ActiveEdges activeEdges,
                                      const uint32_t sliceLo[2], const uint32_t sliceHi[2], int numBuffers,
                                      unsigned long long minBytes, unsigned long long edgeBytes, char *done)
{
    ResidentChoice c;
    for (int b = 0; b < numBuffers && b < 2; b++) {
        const uint32_t lo = sliceLo[b], hi = sliceHi[b];
        if (hi <= lo) continue;
        const size_t i0 = (size_t)(std::lower_bound(bounds, bounds + P + 1, lo) - bounds);
        const size_t i1 = (size_t)(std::lower_bound(bounds, bounds + P + 1, hi) - bounds);
        if (i1 > P || bounds[i0] != lo || bounds[i1] != hi) continue;
        bool taken = false;
        for (size_t i = i0; i < i1; i++) taken = taken || done[i];
        if (taken) {
            c.overlaps++;
            continue;
        }
        unsigned long long active = 0;
        for (size_t i = i0; i < i1; i++) active += activeEdges(i);
        if (active * edgeBytes < minBytes) continue;
        for (size_t i = i0; i < i1; i++) done[i] = 1;
        c.visit[b] = true;
    }
    return c;
}

// Edge cap of a unit of whole id ranges that opens while `remaining` edges
// of whole ranges (this one included) are still to move in the iteration:
// a full buffer while plenty remains, otherwise half of what remains but not
// below minEdges, and never less than the range it opens with.  Units then
// shrink geometrically towards the end of the iteration, so the rounds of
// the last unit, which no transfer overlaps, cover about one range instead
// of a whole buffer; minEdges keeps units whose rounds would take less than
// a launch costs from being split.
constexpr unsigned long long kTailMinEdges = 1ull << 25;
inline unsigned long long whole_unit_cap(unsigned long long bufferEdges, unsigned long long firstRange,
                                         unsigned long long remaining, unsigned long long minEdges = kTailMinEdges)
{
    return std::max(firstRange, std::min(bufferEdges, std::max(remaining / 2, minEdges)));
}

// A whole unit of low-degree rows costs its rounds more per vertex than per
// edge: every round scans all of its vertices.  Its cost is counted as
// edges + kVertexCost * vertices (about three edges of rounds work a vertex
// on the low-degree ranges of a scale-28 R-MAT), and a unit whose vertices
// dominate that cost is cut into pieces: each piece takes half of the cost
// still to come in the iteration (the rest of the unit plus `after`, the
// cost of the whole ranges behind it), not less than minCost, so towards the
// end of an iteration the pieces shrink and the last one, whose rounds no
// transfer overlaps, is short.  Writes the piece ends (ids, ascending, the
// last one b) to `ends`.  np: node pointers (np[v] .. np[v + 1] = row v).
constexpr unsigned long long kVertexCost = 3;
template <class NP, class Out>
inline void split_by_cost(uint32_t a, uint32_t b, const NP *np, unsigned long long after,
                          unsigned long long minCost, Out &ends)
{
    auto cost = [&](uint32_t x, uint32_t y) {
        return (unsigned long long)(np[y] - np[x]) + kVertexCost * (unsigned long long)(y - x);
    };
    while (a < b) {
        const unsigned long long cap = std::max(minCost, (cost(a, b) + after) / 2);
        // the largest end e in (a, b] with cost(a, e) <= cap, at least a + 1
        uint32_t lo = a + 1, hi = b;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo + 1) / 2;
            if (cost(a, mid) <= cap) lo = mid;
            else hi = mid - 1;
        }
        ends.push_back(lo);
        a = lo;
    }
}

}  // namespace subway

#endif  // SUBWAY_SCHEDULE_HPP
