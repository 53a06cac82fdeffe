// argument_parsing.hpp -- the applications' command line (drop-in for
// shared/argument_parsing.cuh:7-36).  Same flags and behaviour: --input
// (required), --output, --source (when allowed), --device (selects the HIP
// device while parsing), --iteration (when allowed); arguments come in
// pairs; a parse error prints the help text and exits with status 0.
// Added flags (all optional): --budget <bytes, K/M/G suffix = KiB/MiB/GiB>,
// --buffers <1|2>, --threshold <PR residual>, --transfer
// <auto|zerocopy|host>, --policy <plain|learned>, --verbose <0|1>.
#ifndef SUBWAY_ARGUMENT_PARSING_HPP
#define SUBWAY_ARGUMENT_PARSING_HPP

#include <string>

#include "globals.hpp"

class ArgumentParser {
public:
    int argc;
    char **argv;

    bool canHaveSource;
    bool canHaveItrs;

    bool hasInput;
    bool hasSourceNode;
    bool hasOutput;
    bool hasDeviceID;
    bool hasNumberOfItrs;
    std::string input;
    int sourceNode;
    std::string output;
    int deviceID;
    int numberOfItrs;

    // additions
    ull budgetBytes;     // 0 = reference sizing
    int numBuffers;
    float threshold;     // PR residual threshold (reference constant 0.01)
    bool hasThreshold;
    int transfer;        // subway::TransferId
    int policy;          // subway::PolicyId
    int verbose;

    ArgumentParser(int argc, char **argv, bool canHaveSource, bool canHaveItrs);
    bool Parse();
    std::string GenerateHelpString();
};

// "1073741824", "1G", "512M", "64K" -> bytes (binary units); 0 on error.
ull parse_size(const std::string &s);

#endif  // SUBWAY_ARGUMENT_PARSING_HPP
