// globals.hpp -- edge layouts, the unreachable sentinel and the integer
// aliases of Subway's public API (shared/globals.hpp:27-51).  The layouts are
// part of the .bcsr/.bwcsr file contract, so they are kept bit-identical:
// OutEdge is one uint32 destination, OutEdgeWeighted is {end, w8} (AoS, 8 B).
#ifndef SUBWAY_GLOBALS_HPP
#define SUBWAY_GLOBALS_HPP

#include <cstdint>
#include <limits>
#include <string>
#include <vector>

typedef unsigned int uint;
typedef unsigned long long ull;

// shared/globals.hpp:27
const unsigned int DIST_INFINITY = std::numeric_limits<unsigned int>::max() - 1;

// shared/globals.hpp:33-35
struct OutEdge {
    uint end;
};

// shared/globals.hpp:37-40
struct OutEdgeWeighted {
    uint end;
    uint w8;
};

// shared/globals.hpp:42-51 (text-format ingestion records)
struct Edge {
    uint source;
    uint end;
};

struct EdgeWeighted {
    uint source;
    uint end;
    uint w8;
};

static_assert(sizeof(OutEdge) == 4, "OutEdge is 4 bytes on disk");
static_assert(sizeof(OutEdgeWeighted) == 8, "OutEdgeWeighted is 8 bytes on disk");

#endif  // SUBWAY_GLOBALS_HPP
