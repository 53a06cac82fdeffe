// Reads "n after minCost" then n + 1 node pointers; prints the piece ends
// split_by_cost gives for the whole id range [0, n).
#include <cstdio>
#include <vector>

#include "schedule.hpp"

int main()
{
    unsigned n = 0;
    unsigned long long after = 0, minCost = 0;
    if (std::scanf("%u %llu %llu", &n, &after, &minCost) != 3) return 1;
    std::vector<unsigned long long> np(n + 1);
    for (unsigned i = 0; i <= n; i++)
        if (std::scanf("%llu", &np[i]) != 1) return 1;
    std::vector<uint32_t> ends;
    subway::split_by_cost(0u, n, np.data(), after, minCost, ends);
    for (uint32_t e : ends) std::printf("%u\n", e);
    return 0;
}
