// gpu_error_check.hpp -- the error convention of the drop-in C++ API
// (shared/gpu_error_check.cuh:8-16): a failing HIP call prints
// "GPUassert: <message> <file> <line>" on stderr and exits with the code.
// Library internals use SUBWAY_CHECK, which throws subway::Error instead so
// that the C ABI can turn failures into return codes.
#ifndef SUBWAY_GPU_ERROR_CHECK_HPP
#define SUBWAY_GPU_ERROR_CHECK_HPP

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#define gpuErrorcheck(ans) { gpuAssert((ans), __FILE__, __LINE__); }
inline void gpuAssert(hipError_t code, const char *file, int line, bool abort = true)
{
    if (code != hipSuccess) {
        fprintf(stderr, "GPUassert: %s %s %d\n", hipGetErrorString(code), file, line);
        if (abort) exit(code);
    }
}

namespace subway {

// Error with a status code: negative values are library errors, positive
// values are hipError_t codes.
struct Error : public std::runtime_error {
    int code;
    Error(int c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

enum Status {
    OK = 0,
    ERR_ARG = -1,       // invalid argument
    ERR_IO = -2,        // missing / unreadable / truncated file
    ERR_FORMAT = -3,    // unsupported extension or malformed content
    ERR_BUDGET = -4,    // device budget too small or above device memory
    ERR_NOMEM = -5,     // host allocation failed
    ERR_INTERNAL = -6,  // invariant violated inside the library
};

inline void check_hip(hipError_t code, const char *expr, const char *file, int line)
{
    if (code != hipSuccess) {
        std::string msg = std::string("GPUassert: ") + hipGetErrorString(code) + " " +
                          file + " " + std::to_string(line) + " (" + expr + ")";
        throw Error((int)code, msg);
    }
}

}  // namespace subway

#define SUBWAY_CHECK(x) ::subway::check_hip((x), #x, __FILE__, __LINE__)

#endif  // SUBWAY_GPU_ERROR_CHECK_HPP
