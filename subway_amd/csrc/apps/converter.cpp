// converter.cpp -- text edge list to binary CSR (reference:
// tools/converter.cpp:20-172).  `converter <file.el|file.wel>` writes
// <file>.bcsr / <file>.bwcsr next to the input, with the reference's layout
// and stable in-source edge order.  A third argument "64" writes the 64-bit
// variant (.bcsr64 / .bwcsr64).  Unlike the reference, per-vertex counters
// are initialised, so small graphs convert correctly.
#include <iostream>
#include <string>

#include "subway/gpu_error_check.hpp"
#include "subway/graph.hpp"

using namespace subway;

int main(int argc, char **argv)
{
    if (argc != 2 && argc != 3) {
        std::cout << "\nThere was an error parsing command line arguments\n";
        return 0;
    }
    const std::string input = argv[1];
    const bool wide = argc == 3 && std::string(argv[2]) == "64";
    const std::string ext = file_extension(input);
    const std::string stem = input.substr(0, input.size() - ext.size());
    try {
        if (ext == "el") {
            auto csr = read_csr_file<OutEdge>(input);
            write_csr_file(*csr, stem + (wide ? "bcsr64" : "bcsr"));
        } else if (ext == "wel") {
            auto csr = read_csr_file<OutEdgeWeighted>(input);
            write_csr_file(*csr, stem + (wide ? "bwcsr64" : "bwcsr"));
        } else {
            std::cout << "\nThe input graph must be an .el or .wel file\n";
            return 0;
        }
    } catch (const Error &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
