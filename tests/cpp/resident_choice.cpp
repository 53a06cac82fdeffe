// Host check of choose_resident_visits (subway_amd/csrc/schedule.hpp): reads
// cases from stdin, one per line:
//   P b0..bP  a0..a(P-1)  lo0 hi0 lo1 hi1  numBuffers minBytes edgeBytes
// and prints "visit0 visit1 overlaps done0..done(P-1)" per case.
#include <cstdio>
#include <vector>

#include "schedule.hpp"

int main()
{
    size_t P;
    while (std::scanf("%zu", &P) == 1) {
        std::vector<uint32_t> bounds(P + 1);
        std::vector<unsigned long long> act(P);
        for (auto &b : bounds) std::scanf("%u", &b);
        for (auto &a : act) std::scanf("%llu", &a);
        uint32_t lo[2], hi[2];
        int nb;
        unsigned long long minBytes, eb;
        std::scanf("%u %u %u %u %d %llu %llu", &lo[0], &hi[0], &lo[1], &hi[1], &nb, &minBytes, &eb);
        std::vector<char> done(P, 0);
        const subway::ResidentChoice c = subway::choose_resident_visits(
            bounds.data(), P, [&](size_t i) { return act[i]; }, lo, hi, nb, minBytes, eb, done.data());
        std::printf("%d %d %u", (int)c.visit[0], (int)c.visit[1], c.overlaps);
        for (char d : done) std::printf(" %d", (int)d);
        std::printf("\n");
    }
    return 0;
}
