// subway_utilities.hpp -- result printing and the --output file (drop-in for
// shared/subway_utilities.hpp:7-14).  PrintResults prints
// "Results of first <n> nodes:\n[0:v0 1:v1 ...]"; SaveResults writes one
// "<i> <value>" line per vertex with default ostream formatting (6
// significant digits for floating point), as shared/subway_utilities.cpp:4-72.
#ifndef SUBWAY_UTILITIES_HPP
#define SUBWAY_UTILITIES_HPP

#include <string>

#include "globals.hpp"

namespace utilities {
void PrintResults(uint *results, uint n);
void PrintResults(float *results, uint n);
void PrintResults(double *results, uint n);
void SaveResults(std::string filepath, uint *results, uint n);
void SaveResults(std::string filepath, float *results, uint n);
void SaveResults(std::string filepath, double *results, uint n);
}  // namespace utilities

#endif  // SUBWAY_UTILITIES_HPP
