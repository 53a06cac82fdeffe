// rmat_gen.cpp -- write a reproducible R-MAT graph as .bcsr/.bwcsr (or the
// 64-bit variants).
//   rmat-gen <scale> <edge_factor> <out.bcsr|out.bwcsr|...> [seed] [sym]
// a=0.57 b=0.19 c=0.19 d=0.05, self-loops dropped, weights U[1,64].
#include <cstdlib>
#include <iostream>
#include <string>

#include "subway_amd.h"

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::cerr << "usage: rmat-gen <scale> <edge_factor> <out.(b|bw)csr[64]> [seed] [symmetrize 0|1]\n";
        return 2;
    }
    int scale = atoi(argv[1]), ef = atoi(argv[2]);
    std::string out = argv[3];
    unsigned long long seed = argc > 4 ? strtoull(argv[4], nullptr, 10) : 42ull;
    int sym = argc > 5 ? atoi(argv[5]) : 0;
    int weighted = out.find(".bwcsr") != std::string::npos;
    subway_graph *g = nullptr;
    int rc = subway_graph_rmat(scale, ef, 0.57, 0.19, 0.19, seed, weighted, sym, 1, 64, &g);
    if (rc == 0) rc = subway_graph_save(g, out.c_str());
    if (rc != 0) {
        std::cerr << subway_last_error() << "\n";
        subway_graph_free(g);
        return 1;
    }
    uint64_t n = 0, m = 0;
    subway_graph_info(g, &n, &m, nullptr, nullptr, nullptr);
    std::cout << "wrote " << out << ": " << n << " vertices, " << m << " edges\n";
    subway_graph_free(g);
    return 0;
}
