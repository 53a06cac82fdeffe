// cli.cpp -- ArgumentParser and result output (reference:
// shared/argument_parsing.cu:5-119, shared/subway_utilities.cpp:4-72).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>

#include "subway/argument_parsing.hpp"
#include "subway/engine.hpp"
#include "subway/subway_utilities.hpp"

ull parse_size(const std::string &s)
{
    if (s.empty()) return 0;
    char *end = nullptr;
    double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || v < 0) return 0;
    double mul = 1;
    if (*end) {
        switch (*end) {
            case 'k': case 'K': mul = 1024.0; break;
            case 'm': case 'M': mul = 1024.0 * 1024.0; break;
            case 'g': case 'G': mul = 1024.0 * 1024.0 * 1024.0; break;
            case 't': case 'T': mul = 1024.0 * 1024.0 * 1024.0 * 1024.0; break;
            default: return 0;
        }
    }
    return (ull)(v * mul);
}

ArgumentParser::ArgumentParser(int argc_, char **argv_, bool canHaveSource_, bool canHaveItrs_)
    : argc(argc_), argv(argv_), canHaveSource(canHaveSource_), canHaveItrs(canHaveItrs_), hasInput(false),
      hasSourceNode(false), hasOutput(false), hasDeviceID(false), hasNumberOfItrs(false), sourceNode(0),
      deviceID(0), numberOfItrs(1), budgetBytes(0), numBuffers(2), threshold(0.01f), hasThreshold(false),
      transfer(subway::TRANSFER_AUTO), policy(subway::POLICY_PLAIN), verbose(0)
{
    Parse();
}

static void fail(const std::string &msg, ArgumentParser &p)
{
    std::cout << msg;
    std::cout << p.GenerateHelpString();
    exit(0);
}

bool ArgumentParser::Parse()
{
    if (argc == 1) {
        std::cout << GenerateHelpString();
        exit(0);
    }
    if (argc == 2 && (!strcmp(argv[1], "--help") || !strcmp(argv[1], "-help") || !strcmp(argv[1], "--h") ||
                      !strcmp(argv[1], "-h"))) {
        std::cout << GenerateHelpString();
        exit(0);
    }
    if (argc % 2 == 0) fail("\nThere was an error parsing command line arguments\n", *this);
    for (int i = 1; i < argc - 1; i += 2) {
        const char *k = argv[i];
        std::string v = argv[i + 1];
        if (!strcmp(k, "--input")) {
            input = v;
            hasInput = true;
        } else if (!strcmp(k, "--output")) {
            output = v;
            hasOutput = true;
        } else if (!strcmp(k, "--source") && canHaveSource) {
            sourceNode = atoi(v.c_str());
            hasSourceNode = true;
        } else if (!strcmp(k, "--device")) {
            deviceID = atoi(v.c_str());
            hasDeviceID = true;
            (void)hipSetDevice(deviceID);
        } else if (!strcmp(k, "--iteration") && canHaveItrs) {
            numberOfItrs = atoi(v.c_str());
            hasNumberOfItrs = true;
        } else if (!strcmp(k, "--budget")) {
            budgetBytes = parse_size(v);
            if (!budgetBytes) fail("\nInvalid --budget <" + v + ">\n", *this);
        } else if (!strcmp(k, "--buffers")) {
            numBuffers = atoi(v.c_str());
            if (numBuffers != 1 && numBuffers != 2) fail("\n--buffers must be 1 or 2\n", *this);
        } else if (!strcmp(k, "--threshold")) {
            threshold = (float)atof(v.c_str());
            hasThreshold = true;
        } else if (!strcmp(k, "--transfer")) {
            if (v == "auto") transfer = subway::TRANSFER_AUTO;
            else if (v == "zerocopy") transfer = subway::TRANSFER_ZEROCOPY;
            else if (v == "host") transfer = subway::TRANSFER_HOST;
            else fail("\nInvalid --transfer <" + v + ">\n", *this);
        } else if (!strcmp(k, "--policy")) {
            if (v == "plain") policy = subway::POLICY_PLAIN;
            else if (v == "learned") policy = subway::POLICY_LEARNED;
            else fail("\nInvalid --policy <" + v + ">\n", *this);
        } else if (!strcmp(k, "--verbose")) {
            verbose = atoi(v.c_str());
        } else {
            fail(std::string("\nThere was an error parsing command line argument <") + k + ">\n", *this);
        }
    }
    if (!hasInput) fail("\nInput graph file argument is required.\n", *this);
    return true;
}

std::string ArgumentParser::GenerateHelpString()
{
    std::string str = "\nRequired arguments:";
    str += "\n    [--input]: Input graph file. E.g., --input FacebookGraph.txt";
    str += "\nOptional arguments";
    if (canHaveSource) str += "\n    [--source]:  Begins from the source (Default: 0). E.g., --source 10";
    str += "\n    [--output]: Output file for results. E.g., --output results.txt";
    str += "\n    [--device]: Select GPU device (default: 0). E.g., --device 1";
    if (canHaveItrs) str += "\n    [--iteration]: Number of iterations (default: 1). E.g., --iterations 10";
    str += "\n    [--budget]: Device edge budget in bytes, K/M/G suffixes (default: 90% of device memory)";
    str += "\n    [--buffers]: Device edge buffers, 1 or 2 (default: 2, overlaps transfer and compute)";
    str += "\n    [--threshold]: PageRank residual threshold (default: 0.01)";
    str += "\n    [--transfer]: auto | zerocopy | host (default: auto)";
    str += "\n    [--policy]: plain | learned (default: plain)";
    str += "\n    [--verbose]: 1 prints one line per partition (default: 0)";
    str += "\n\n";
    return str;
}

namespace utilities {

template <class T>
static void print_results(T *results, uint n)
{
    std::cout << "Results of first " << n << " nodes:\n[";
    for (uint i = 0; i < n; i++) {
        if (i > 0) std::cout << " ";
        std::cout << i << ":" << results[i];
    }
    std::cout << "]\n";
}

void PrintResults(uint *results, uint n) { print_results(results, n); }
void PrintResults(float *results, uint n) { print_results(results, n); }
void PrintResults(double *results, uint n) { print_results(results, n); }

template <class T>
static void save_results(const std::string &filepath, T *results, uint n)
{
    std::ofstream outfile(filepath);
    for (uint i = 0; i < n; i++) outfile << i << " " << results[i] << "\n";
}

void SaveResults(std::string filepath, uint *results, uint n)
{
    std::cout << "Saving the results into the following file:\n";
    std::cout << ">> " << filepath << std::endl;
    save_results(filepath, results, n);
    std::cout << "Done saving.\n";
}

void SaveResults(std::string filepath, float *results, uint n)
{
    std::cout << "Saving the results into " << filepath << " ...... " << std::flush;
    save_results(filepath, results, n);
    std::cout << " Completed.\n";
}

void SaveResults(std::string filepath, double *results, uint n)
{
    std::cout << "Saving the results into " << filepath << " ...... " << std::flush;
    save_results(filepath, results, n);
    std::cout << " Completed.\n";
}

}  // namespace utilities
