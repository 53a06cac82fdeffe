// Reads "buffer first remaining min" lines, prints whole_unit_cap for each.
#include <cstdio>

#include "schedule.hpp"

int main()
{
    unsigned long long b, f, r, m;
    while (std::scanf("%llu %llu %llu %llu", &b, &f, &r, &m) == 4)
        std::printf("%llu\n", subway::whole_unit_cap(b, f, r, m));
    return 0;
}
