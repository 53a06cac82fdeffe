// noalloc.hip -- subway_generate_noalloc: the subgraph generator of
// subway_generate over caller-owned device arrays, with its scratch (the scan
// control block and the per-chunk count/degree aggregates) in a file-scope
// device array, so a call allocates nothing.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "kernels.hpp"
#include "subway/gpu_error_check.hpp"
#include "subway_amd.h"

namespace subway {
void set_last_error(const char *msg);   // capi.cpp

namespace {

// Largest vertex count a call accepts (vertex ids are 32-bit); one count and
// one degree-sum aggregate of 8 bytes per 16384-vertex chunk of the
// generator, after a 256-byte control block.
constexpr unsigned long long kMaxVertices = 0xFFFFFFFFull;
constexpr unsigned long long kMaxChunks = (kMaxVertices + GEN_CHUNK - 1) / GEN_CHUNK;
constexpr size_t kCtlBytes = 256;
constexpr size_t kScratchBytes = kCtlBytes + 2 * kMaxChunks * sizeof(unsigned long long);
static_assert(sizeof(ScanCtl) <= kCtlBytes, "the control block fits its slot");

// Each GPU holds its own copy until the process ends.  Host code never uses
// this name as an address: it uses the per-device address from
// hipGetSymbolAddress below.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_addr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device.
void *scratch_address()
{
    int dev = 0;
    SUBWAY_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(ERR_ARG, "device index beyond the no-allocation scratch table");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_addr[dev]) {
        void *p = nullptr;
        SUBWAY_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(ERR_INTERNAL, "no device address for the generator scratch");
        g_addr[dev] = p;
    }
    return g_addr[dev];
}

}  // namespace
}  // namespace subway

using namespace subway;

extern "C" int32_t subway_generate_noalloc(uint64_t n, const uint8_t *label1, const uint8_t *label2,
                                           const uint32_t *out_degree, uint32_t *active_nodes,
                                           uint64_t *active_nodes_pointer, uint64_t *num_active)
{
    try {
        if (!n || !label1 || !label2 || !out_degree || !active_nodes || !active_nodes_pointer || !num_active)
            throw Error(ERR_ARG, "null argument or empty graph");
        if (n > kMaxVertices)
            throw Error(ERR_ARG, "subway_generate_noalloc takes at most " + std::to_string(kMaxVertices) +
                                     " vertices (its fixed scratch holds " + std::to_string(kMaxChunks) +
                                     " chunk aggregates); got " + std::to_string(n));
        char *base = static_cast<char *>(scratch_address());
        ScanScratch s;
        s.ctl = reinterpret_cast<ScanCtl *>(base);
        s.td.flag = reinterpret_cast<unsigned long long *>(base + kCtlBytes);
        s.td.inc = s.td.flag + kMaxChunks;
        s.maxTiles = kMaxChunks;
        // the control block starts at zero (error word, histogram); a fill
        // kernel, not a memset, so the call can sit in a captured graph
        launch_fill_u32(reinterpret_cast<uint32_t *>(s.ctl), 0u, kCtlBytes / sizeof(uint32_t), nullptr);
        launch_generate_labels(n, label1, label2, out_degree, active_nodes,
                               reinterpret_cast<ull *>(active_nodes_pointer), s, nullptr);
        unsigned long long totals[3] = {0, 0, 0};
        SUBWAY_CHECK(hipMemcpy(totals, &s.ctl->count, sizeof(totals), hipMemcpyDeviceToHost));
        if (totals[2]) throw Error(ERR_INTERNAL, "generator: a device wait gave up");
        *num_active = totals[0];
        return SUBWAY_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code ? e.code : SUBWAY_ERR_INTERNAL;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return SUBWAY_ERR_INTERNAL;
    }
}
