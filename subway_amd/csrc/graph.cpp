// graph.cpp -- CSR storage, the on-disk formats and the Graph / GraphPR
// classes (reference: shared/graph.cu, tools/converter.cpp).
//
// Formats (extension decides, as in shared/graph.cu:36-52):
//   .bcsr / .bwcsr     [u32 n][u32 m][u32 nodePointer x n][E x m]
//                      (shared/graph.cu:39-51, tools/converter.cpp:88-93)
//   .bcsr64 / .bwcsr64 [u64 n][u64 m][u64 nodePointer x n][E x m]
//                      64-bit variant for m >= 2^32 (same layout, wider ints)
//   .el / .wel         text, one "source end [w8]" per line; n = max id + 1;
//                      stable counting sort by source (shared/graph.cu:53-170)
// OutEdge is {u32 end}; OutEdgeWeighted is {u32 end, u32 w8}.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <thread>
#include <type_traits>
#include <vector>

#include "subway/globals.hpp"
#include "subway/gpu_error_check.hpp"
#include "subway/graph.hpp"
#include "host_util.hpp"

namespace subway {

template <class E>
constexpr bool is_weighted_edge() { return sizeof(E) == 8; }

std::string file_extension(const std::string &fileName)
{
    size_t p = fileName.find_last_of('.');
    if (p != std::string::npos) return fileName.substr(p + 1);
    return "";
}

template <class E>
HostCsr<E>::~HostCsr()
{
    delete[] ptr;
    delete[] degree;
    if (edges) {
        if (pinned) (void)hipHostFree(edges);
        else std::free(edges);
    }
    if (packed) (void)hipHostFree(packed);
    if (splitPinned) {
        if (splitDst) (void)hipHostFree(splitDst);
        if (splitW) (void)hipHostFree(splitW);
    } else {
        std::free(splitDst);
        std::free(splitW);
    }
    delete[] perm;
    delete[] iperm;
    if (blk) (void)hipHostFree(blk);
    delete[] blkPtr;
    if (tailDst) (void)hipHostFree(tailDst);
    if (tailW) (void)hipHostFree(tailW);
}

template <class E>
uint32_t HostCsr<E>::ensurePacked()
{
    if constexpr (!std::is_same<E, OutEdgeWeighted>::value) {
        return 0;
    } else {
        std::lock_guard<std::mutex> lock(packMu_);
        if (packTried_) return packShift;
        packTried_ = true;
        if (!pinned || m == 0) return 0;
        uint32_t db = 1;
        while (db < 32 && (1ull << db) < n) db++;   // ids 0 .. n-1 fit db bits
        if (db >= 32) return 0;
        const int T = host_threads();
        std::vector<uint32_t> mx((size_t)T, 0), mn((size_t)T, 0xFFFFFFFFu);
        parallel_for(m, [&](uint64_t b, uint64_t e, int t) {
            uint32_t x = 0, y = 0xFFFFFFFFu;
            for (uint64_t i = b; i < e; i++) {
                x = std::max(x, edges[i].w8);
                y = std::min(y, edges[i].w8);
            }
            mx[(size_t)t] = std::max(mx[(size_t)t], x);
            mn[(size_t)t] = std::min(mn[(size_t)t], y);
        }, T);
        const uint32_t maxW = *std::max_element(mx.begin(), mx.end());
        uint32_t base = *std::min_element(mn.begin(), mn.end());
        if (base >> 27) base = 0;                         // the decode word holds a base below 2^27
        if ((uint64_t)(maxW - base) >> (32 - db)) return 0;   // the weight span needs more than 32 - db bits
        uint32_t *p = nullptr;
        if (hipHostMalloc((void **)&p, (size_t)m * sizeof(uint32_t), hipHostMallocMapped | hipHostMallocPortable) !=
            hipSuccess) {
            (void)hipGetLastError();
            return 0;   // no room for the packed copy: transfers keep the 8-byte edges
        }
        parallel_for(m, [&](uint64_t b, uint64_t e, int) {
            for (uint64_t i = b; i < e; i++) p[i] = edges[i].end | ((edges[i].w8 - base) << db);
        }, T);
        packed = p;
        SUBWAY_CHECK(hipHostGetDevicePointer((void **)&d_packed, packed, 0));
        packBase = base;
        packShift = db;
        return db;
    }
}

template <class E>
bool HostCsr<E>::ensureSplit()
{
    if constexpr (!std::is_same<E, OutEdgeWeighted>::value) {
        return false;
    } else {
        std::lock_guard<std::mutex> lock(packMu_);
        if (splitTried_) return splitDst != nullptr;
        splitTried_ = true;
        if (!pinned || m == 0) return false;
        const int T = host_threads();
        std::vector<uint32_t> mx((size_t)T, 0);
        parallel_for(m, [&](uint64_t b, uint64_t e, int t) {
            uint32_t x = 0;
            for (uint64_t i = b; i < e; i++) x = std::max(x, edges[i].w8);
            mx[(size_t)t] = std::max(mx[(size_t)t], x);
        }, T);
        if (*std::max_element(mx.begin(), mx.end()) > 0xFFu) return false;
        uint32_t *d = nullptr;
        uint8_t *w = nullptr;
        const unsigned flags = hipHostMallocMapped | hipHostMallocPortable;
        hipError_t e = hipHostMalloc((void **)&d, (size_t)m * sizeof(uint32_t), flags);
        if (e == hipSuccess) e = hipHostMalloc((void **)&w, (size_t)m, flags);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            if (d) (void)hipHostFree(d);
            // no room for the split copy: transfers keep the 8-byte edges
            std::fprintf(stderr, "subway: pinned split copy of %llu edges not allocated (%s); 8-byte transfers\n",
                         (unsigned long long)m, hipGetErrorString(e));
            return false;
        }
        parallel_for(m, [&](uint64_t b, uint64_t e, int) {
            for (uint64_t i = b; i < e; i++) {
                d[i] = edges[i].end;
                w[i] = (uint8_t)edges[i].w8;
            }
        }, T);
        splitDst = d;
        splitW = w;
        SUBWAY_CHECK(hipHostGetDevicePointer((void **)&d_splitDst, splitDst, 0));
        SUBWAY_CHECK(hipHostGetDevicePointer((void **)&d_splitW, splitW, 0));
        return true;
    }
}

template <class E>
void HostCsr<E>::demoteSplit()
{
    std::lock_guard<std::mutex> lock(packMu_);
    if (!splitDst || !splitPinned) return;
    const size_t mm = (size_t)(m ? m : 1);
    uint32_t *d = static_cast<uint32_t *>(std::malloc(mm * sizeof(uint32_t)));
    uint8_t *w = static_cast<uint8_t *>(std::malloc(mm));
    if (!d || !w) {
        std::free(d);
        std::free(w);
        throw Error(ERR_NOMEM, "host allocation for the split edge arrays failed");
    }
    parallel_for(m, [&](uint64_t b, uint64_t e, int) {
        std::memcpy(d + b, splitDst + b, (e - b) * sizeof(uint32_t));
        std::memcpy(w + b, splitW + b, e - b);
    });
    (void)hipHostFree(splitDst);
    (void)hipHostFree(splitW);
    splitDst = d;
    splitW = w;
    d_splitDst = nullptr;
    d_splitW = nullptr;
    splitPinned = false;
}

template <class E>
bool HostCsr<E>::allocateSplit()
{
    std::lock_guard<std::mutex> lock(packMu_);
    splitTried_ = true;
    const unsigned flags = hipHostMallocMapped | hipHostMallocPortable;
    uint32_t *d = nullptr;
    uint8_t *w = nullptr;
    hipError_t e = hipHostMalloc((void **)&d, (size_t)(m ? m : 1) * sizeof(uint32_t), flags);
    if (e == hipSuccess) e = hipHostMalloc((void **)&w, (size_t)(m ? m : 1), flags);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        if (d) (void)hipHostFree(d);
        return false;
    }
    splitDst = d;
    splitW = w;
    SUBWAY_CHECK(hipHostGetDevicePointer((void **)&d_splitDst, splitDst, 0));
    SUBWAY_CHECK(hipHostGetDevicePointer((void **)&d_splitW, splitW, 0));
    return true;
}

template <class E>
void HostCsr<E>::allocate(ull n_, ull m_, bool pinEdges, bool withEdges)
{
    if (n_ == 0) throw Error(ERR_FORMAT, "graph has no vertices");
    if (n_ > 0xFFFFFFFFull) throw Error(ERR_FORMAT, "vertex ids are 32-bit; graph has too many vertices");
    n = n_;
    m = m_;
    try {
        ptr = new ull[n + 1];
        degree = new uint[n];
    } catch (const std::bad_alloc &) {
        throw Error(ERR_NOMEM, "host allocation of the CSR offsets failed");
    }
    size_t bytes = (size_t)(m ? m : 1) * sizeof(E);
    if (!withEdges) {
        edges = nullptr;
        pinned = false;
        return;
    }
    if (!pinEdges) {
        edges = static_cast<E *>(std::malloc(bytes));
        if (!edges) throw Error(ERR_NOMEM, "host allocation of the edge list failed");
        pinned = false;
        return;
    }
    hipError_t e = hipHostMalloc((void **)&edges, bytes, hipHostMallocMapped | hipHostMallocPortable);
    if (e != hipSuccess) {
        edges = nullptr;
        int devices = 0;
        if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) {
            // No GPU in this process: the CSR can still be built, read and
            // written (file conversion, generation); running the pipeline on
            // it is refused because nothing is mapped for the device.
            (void)hipGetLastError();
            edges = static_cast<E *>(std::malloc(bytes));
            if (!edges) throw Error(ERR_NOMEM, "host allocation of the edge list failed");
            pinned = false;
            return;
        }
        throw Error(ERR_NOMEM, "pinned allocation of " + std::to_string(bytes) +
                                   " bytes for the edge list failed: " + hipGetErrorString(e));
    }
    pinned = true;
}

template <class E>
void HostCsr<E>::finalize()
{
    ptr[n] = m;
    uint mx = 0;
    for (ull i = 0; i < n; i++) {
        if (ptr[i + 1] < ptr[i]) throw Error(ERR_FORMAT, "node pointers are not non-decreasing");
        ull d = ptr[i + 1] - ptr[i];
        if (d > 0xFFFFFFFFull) throw Error(ERR_FORMAT, "a vertex has 2^32 or more edges");
        degree[i] = (uint)d;
        mx = std::max(mx, (uint)d);
    }
    maxDegree = mx;
    large = large || (m > 0xFFFFFFFFull);
    if (pinned) SUBWAY_CHECK(hipHostGetDevicePointer((void **)&d_edges, edges, 0));
}

// A weighted edge list read into ordinary memory (it may exceed what a
// process can pin): when its weights fit a byte and ids + weight span do not
// pack into one word (or the caller asks for it), the pinned memory becomes
// the split arrays (5 bytes per edge) and the 8-byte list stays in ordinary
// memory; otherwise the list moves into pinned memory as every other graph's.
template <class E>
void HostCsr<E>::settleEdges(bool forceSplit)
{
    if (pinned) return;
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) {
        (void)hipGetLastError();
        return;   // no GPU in this process: nothing is mapped (conversion, inspection)
    }
    bool split = false;
    if constexpr (std::is_same<E, OutEdgeWeighted>::value) {
        const int T = host_threads();
        std::vector<uint32_t> mx((size_t)T, 0), mn((size_t)T, 0xFFFFFFFFu);
        parallel_for(m, [&](uint64_t b, uint64_t e, int t) {
            uint32_t x = 0, y = 0xFFFFFFFFu;
            for (uint64_t i = b; i < e; i++) {
                x = std::max(x, edges[i].w8);
                y = std::min(y, edges[i].w8);
            }
            mx[(size_t)t] = std::max(mx[(size_t)t], x);
            mn[(size_t)t] = std::min(mn[(size_t)t], y);
        }, T);
        const uint32_t maxW = m ? *std::max_element(mx.begin(), mx.end()) : 0;
        uint32_t base = m ? *std::min_element(mn.begin(), mn.end()) : 0;
        if (base >> 27) base = 0;
        uint32_t db = 1;
        while (db < 32 && (1ull << db) < n) db++;
        const bool packs = db < 32 && ((uint64_t)(maxW - base) >> (32 - db)) == 0;
        if (forceSplit && maxW > 0xFFu) throw Error(ERR_ARG, "the split layout needs weights that fit a byte");
        split = forceSplit || (!packs && maxW <= 0xFFu);
        if (split) {
            if (!allocateSplit()) throw Error(ERR_NOMEM, "pinned allocation of the split edge arrays failed");
            parallel_for(m, [&](uint64_t b, uint64_t e, int) {
                for (uint64_t i = b; i < e; i++) {
                    splitDst[i] = edges[i].end;
                    splitW[i] = (uint8_t)edges[i].w8;
                }
            }, T);
            return;
        }
    }
    const size_t bytes = (size_t)(m ? m : 1) * sizeof(E);
    E *p = nullptr;
    const hipError_t e = hipHostMalloc((void **)&p, bytes, hipHostMallocMapped | hipHostMallocPortable);
    if (e != hipSuccess)
        throw Error(ERR_NOMEM, "pinned allocation of " + std::to_string(bytes) +
                                   " bytes for the edge list failed: " + hipGetErrorString(e));
    parallel_for(m, [&](uint64_t b, uint64_t e2, int) { std::memcpy(p + b, edges + b, (e2 - b) * sizeof(E)); });
    std::free(edges);
    edges = p;
    pinned = true;
}

template <class E>
void HostCsr<E>::adoptPacked(uint32_t *p, uint32_t shift, uint32_t base)
{
    std::lock_guard<std::mutex> lock(packMu_);
    packTried_ = true;
    packed = p;
    packShift = shift;
    packBase = base;
    SUBWAY_CHECK(hipHostGetDevicePointer((void **)&d_packed, packed, 0));
}

// ------------------------------------------------------------ binary I/O
static void write_exact(FILE *f, const void *src, size_t bytes, const std::string &path)
{
    const char *p = (const char *)src;
    while (bytes) {
        size_t chunk = std::min(bytes, (size_t)1 << 30);
        if (fwrite(p, 1, chunk, f) != chunk) throw Error(ERR_IO, "write failed: " + path);
        p += chunk;
        bytes -= chunk;
    }
}

// Positioned reads of [off, off + bytes) of the file into dst by host
// threads, 64 MiB blocks claimed in order (several reads in flight keep a
// fast device or the page cache busy; one stream of fread does not).
static void pread_all(int fd, void *dst, ull bytes, ull off, const std::string &path)
{
    constexpr ull BLK = 64ull << 20;
    const ull nb = (bytes + BLK - 1) / BLK;
    std::atomic<ull> next{0};
    std::atomic<int> bad{0};
    const int T = (int)std::max<ull>(1, std::min<ull>((ull)std::min(host_threads(), 16), nb));
    auto work = [&] {
        for (ull b; (b = next.fetch_add(1)) < nb && !bad.load();) {
            char *p = (char *)dst + b * BLK;
            ull left = std::min(BLK, bytes - b * BLK), at = off + b * BLK;
            while (left) {
                const ssize_t got = ::pread(fd, p, (size_t)std::min<ull>(left, 1ull << 30), (off_t)at);
                if (got <= 0) {
                    bad.store(got < 0 ? errno : -1);
                    break;
                }
                p += got;
                at += (ull)got;
                left -= (ull)got;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < T; t++) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    if (bad.load()) throw Error(ERR_IO, "read of the graph file failed: " + path);
}

// A weighted edge region read straight into the pinned split arrays, block
// by block through per-thread staging: no 8-byte list is ever held.  Returns
// false (nothing kept) when a weight does not fit a byte.
static bool read_split(int fd, HostCsr<OutEdgeWeighted> &csr, ull off, const std::string &path, uint32_t &minW,
                       uint32_t &maxW)
{
    if (!csr.allocateSplit()) throw Error(ERR_NOMEM, "pinned allocation of the split edge arrays failed");
    constexpr ull BLK = 8ull << 20;   // edges per block (64 MiB)
    const ull m = csr.m, nb = (m + BLK - 1) / BLK;
    std::atomic<ull> next{0};
    std::atomic<int> bad{0};
    std::atomic<uint32_t> mx{0}, mn{0xFFFFFFFFu}, maxEnd{0};
    const int T = (int)std::max<ull>(1, std::min<ull>((ull)std::min(host_threads(), 16), nb));
    auto work = [&] {
        std::vector<OutEdgeWeighted> buf((size_t)std::min(BLK, m ? m : 1));
        uint32_t lmx = 0, lmn = 0xFFFFFFFFu, lend = 0;
        for (ull b; (b = next.fetch_add(1)) < nb && !bad.load();) {
            const ull e0 = b * BLK, cnt = std::min(BLK, m - e0);
            char *p = (char *)buf.data();
            ull left = cnt * sizeof(OutEdgeWeighted), at = off + e0 * sizeof(OutEdgeWeighted);
            while (left) {
                const ssize_t got = ::pread(fd, p, (size_t)left, (off_t)at);
                if (got <= 0) {
                    bad.store(1);
                    break;
                }
                p += got;
                at += (ull)got;
                left -= (ull)got;
            }
            for (ull i = 0; i < cnt; i++) {
                const OutEdgeWeighted x = buf[(size_t)i];
                csr.splitDst[e0 + i] = x.end;
                csr.splitW[e0 + i] = (uint8_t)x.w8;
                lmx = std::max(lmx, x.w8);
                lmn = std::min(lmn, x.w8);
                lend = std::max(lend, x.end);
            }
        }
        for (uint32_t c = mx.load(); lmx > c && !mx.compare_exchange_weak(c, lmx);) {}
        for (uint32_t c = mn.load(); lmn < c && !mn.compare_exchange_weak(c, lmn);) {}
        for (uint32_t c = maxEnd.load(); lend > c && !maxEnd.compare_exchange_weak(c, lend);) {}
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < T; t++) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    if (bad.load()) throw Error(ERR_IO, "truncated graph file: " + path);
    if (m && maxEnd.load() >= csr.n) throw Error(ERR_FORMAT, "edge endpoint out of range: " + path);
    maxW = m ? mx.load() : 0;
    minW = m ? mn.load() : 0;
    return maxW <= 0xFFu;
}

template <class E>
static void read_binary(HostCsr<E> &csr, const std::string &path, bool wide, bool splitLayout)
{
    const int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) throw Error(ERR_IO, "cannot open graph file: " + path + " (" + strerror(errno) + ")");
    struct Closer { int fd; ~Closer() { ::close(fd); } } closer{fd};
    struct stat sb;
    if (::fstat(fd, &sb) != 0) throw Error(ERR_IO, "cannot stat graph file: " + path);
    const ull size = (ull)sb.st_size;
    ull n, m;
    const size_t hdr = wide ? 16 : 8;
    if (size < hdr) throw Error(ERR_IO, "truncated graph file header: " + path);
    if (wide) {
        uint64_t h[2];
        pread_all(fd, h, 16, 0, path);
        n = h[0];
        m = h[1];
    } else {
        uint32_t h[2];
        pread_all(fd, h, 8, 0, path);
        n = h[0];
        m = h[1];
    }
    const ull ptrBytes = n * (wide ? 8ull : 4ull);
    const ull expect = hdr + ptrBytes + m * sizeof(E);
    if (size != expect)
        throw Error(ERR_FORMAT, "graph file size " + std::to_string(size) + " does not match header (n=" +
                                    std::to_string(n) + ", m=" + std::to_string(m) + ", expected " +
                                    std::to_string(expect) + " bytes): " + path);
    const ull edgeOff = hdr + ptrBytes;
    // weighted lists of 16 GiB or more (or a requested split layout) go
    // straight into the pinned split arrays when their weights fit a byte
    bool settled = false;
    if constexpr (is_weighted_edge<E>()) {
        int devices = 0;
        const bool gpu = hipGetDeviceCount(&devices) == hipSuccess && devices > 0;
        if (!gpu) (void)hipGetLastError();
        if (gpu && (splitLayout || m * sizeof(E) >= (1ull << 34))) {
            csr.allocate(n, m, false, /*withEdges=*/false);
            uint32_t minW = 0, maxW = 0;
            if (read_split(fd, csr, edgeOff, path, minW, maxW)) {
                settled = true;
                // ids and weight span in one word: the transfers use a packed
                // copy built from the split arrays (4 bytes per edge)
                uint32_t db = 1;
                while (db < 32 && (1ull << db) < n) db++;
                uint32_t base = minW >> 27 ? 0 : minW;
                if (!splitLayout && db < 32 && ((uint64_t)(maxW - base) >> (32 - db)) == 0 && m) {
                    uint32_t *p = nullptr;
                    if (hipHostMalloc((void **)&p, (size_t)m * sizeof(uint32_t),
                                      hipHostMallocMapped | hipHostMallocPortable) == hipSuccess) {
                        parallel_for(m, [&](uint64_t b, uint64_t e, int) {
                            for (uint64_t i = b; i < e; i++)
                                p[i] = csr.splitDst[i] | (((uint32_t)csr.splitW[i] - base) << db);
                        });
                        csr.adoptPacked(p, db, base);
                    } else {
                        (void)hipGetLastError();
                    }
                }
            } else if (splitLayout) {
                throw Error(ERR_ARG, "the split layout needs weights that fit a byte");
            } else {
                // weights wider than a byte: the 8-byte list after all
                if (csr.splitDst) (void)hipHostFree(csr.splitDst);
                if (csr.splitW) (void)hipHostFree(csr.splitW);
                csr.splitDst = csr.d_splitDst = nullptr;
                csr.splitW = csr.d_splitW = nullptr;
                delete[] csr.ptr;
                delete[] csr.degree;
                csr.ptr = nullptr;
                csr.degree = nullptr;
            }
        }
    }
    if (!settled) {
        if (!csr.ptr) csr.allocate(n, m, true);
        pread_all(fd, csr.edges, m * sizeof(E), edgeOff, path);
        std::atomic<uint32_t> maxEnd{0};
        parallel_for(m, [&](uint64_t b, uint64_t e, int) {
            uint32_t x = 0;
            for (uint64_t i = b; i < e; i++) x = std::max(x, csr.edges[i].end);
            for (uint32_t c = maxEnd.load(); x > c && !maxEnd.compare_exchange_weak(c, x);) {}
        });
        if (m && maxEnd.load() >= n) throw Error(ERR_FORMAT, "edge endpoint out of range: " + path);
    }
    csr.large = wide;
    if (wide) {
        pread_all(fd, csr.ptr, n * 8, hdr, path);
    } else {
        std::vector<uint32_t> p32(n);
        pread_all(fd, p32.data(), n * 4, hdr, path);
        parallel_for(n, [&](uint64_t b, uint64_t e, int) {
            for (uint64_t i = b; i < e; i++) csr.ptr[i] = p32[i];
        });
    }
    for (ull i = 0; i < n; i++)
        if (csr.ptr[i] > m) throw Error(ERR_FORMAT, "node pointer beyond edge count: " + path);
    csr.finalize();
}

// ------------------------------------------------------------- text I/O
static bool parse_uint(const char *&p, const char *end, uint64_t &out)
{
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == ',')) p++;
    if (p >= end || *p < '0' || *p > '9') return false;
    uint64_t v = 0;
    while (p < end && *p >= '0' && *p <= '9') {
        v = v * 10 + (uint64_t)(*p - '0');
        if (v > 0xFFFFFFFFull) return false;
        p++;
    }
    out = v;
    return true;
}

template <class E>
static void read_text(HostCsr<E> &csr, const std::string &path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) throw Error(ERR_IO, "cannot open graph file: " + path);
    std::vector<uint32_t> src, dst, wt;
    std::string line;
    uint64_t mx = 0;
    bool any = false;
    unsigned long long lineNo = 0;
    while (std::getline(in, line)) {
        lineNo++;
        const char *p = line.data(), *end = p + line.size();
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) p++;
        if (p == end || *p == '#' || *p == '%') continue;  // blank or comment line
        uint64_t s, d, w = 0;
        if (!parse_uint(p, end, s) || !parse_uint(p, end, d))
            throw Error(ERR_FORMAT, path + ":" + std::to_string(lineNo) + ": expected 'source end'");
        if (is_weighted_edge<E>() && !parse_uint(p, end, w))
            throw Error(ERR_FORMAT, path + ":" + std::to_string(lineNo) + ": expected a weight");
        src.push_back((uint32_t)s);
        dst.push_back((uint32_t)d);
        if (is_weighted_edge<E>()) wt.push_back((uint32_t)w);
        mx = std::max(mx, std::max(s, d));
        any = true;
    }
    if (!any) throw Error(ERR_FORMAT, "graph file has no edges: " + path);
    ull n = mx + 1, m = src.size();
    csr.allocate(n, m);
    std::vector<ull> cursor(n, 0);
    for (ull e = 0; e < m; e++) cursor[src[e]]++;
    ull run = 0;
    for (ull v = 0; v < n; v++) {
        csr.ptr[v] = run;
        ull d = cursor[v];
        cursor[v] = run;
        run += d;
    }
    for (ull e = 0; e < m; e++) {
        ull at = cursor[src[e]]++;
        csr.edges[at].end = dst[e];
        if constexpr (sizeof(E) == 8) reinterpret_cast<uint32_t *>(&csr.edges[at])[1] = wt[e];
    }
    csr.finalize();
}

template <class E>
std::shared_ptr<HostCsr<E>> read_csr_file(const std::string &path, bool splitLayout)
{
    auto csr = std::make_shared<HostCsr<E>>();
    std::string ext = file_extension(path);
    bool weighted = is_weighted_edge<E>();
    if (ext == "bcsr" || ext == "bcsr64") {
        if (weighted) throw Error(ERR_FORMAT, "a weighted graph needs a .bwcsr file, got " + path);
        read_binary(*csr, path, ext == "bcsr64", false);
    } else if (ext == "bwcsr" || ext == "bwcsr64") {
        if (!weighted) throw Error(ERR_FORMAT, "an unweighted graph needs a .bcsr file, got " + path);
        read_binary(*csr, path, ext == "bwcsr64", splitLayout);
    } else if (ext == "el" || ext == "wel") {
        if (splitLayout) throw Error(ERR_ARG, "the split layout is read from .bwcsr files only");
        read_text(*csr, path);
    } else {
        throw Error(ERR_FORMAT, "The graph format is not supported: " + path);
    }
    return csr;
}

template <class E>
void write_csr_file(const HostCsr<E> &csr, const std::string &path)
{
    std::string ext = file_extension(path);
    bool wide;
    if (ext == "bcsr" || ext == "bwcsr") wide = false;
    else if (ext == "bcsr64" || ext == "bwcsr64") wide = true;
    else throw Error(ERR_FORMAT, "output must be .bcsr/.bwcsr or the 64-bit variants: " + path);
    bool wantWeighted = (ext == "bwcsr" || ext == "bwcsr64");
    if (wantWeighted != is_weighted_edge<E>())
        throw Error(ERR_FORMAT, "file extension does not match the edge type: " + path);
    if (!wide && (csr.m > 0xFFFFFFFFull || csr.n > 0xFFFFFFFFull))
        throw Error(ERR_FORMAT, "graph too large for the 32-bit header; use the 64-bit variant");
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) throw Error(ERR_IO, "cannot create " + path);
    struct Closer { FILE *f; ~Closer() { fclose(f); } } closer{f};
    if (wide) {
        uint64_t h[2] = {csr.n, csr.m};
        write_exact(f, h, 16, path);
        write_exact(f, csr.ptr, csr.n * 8, path);
    } else {
        uint32_t h[2] = {(uint32_t)csr.n, (uint32_t)csr.m};
        write_exact(f, h, 8, path);
        std::vector<uint32_t> p32(csr.n);
        for (ull i = 0; i < csr.n; i++) p32[i] = (uint32_t)csr.ptr[i];
        write_exact(f, p32.data(), csr.n * 4, path);
    }
    if (csr.edges) {
        write_exact(f, csr.edges, csr.m * sizeof(E), path);
        return;
    }
    // split arrays only: the file layout is built in blocks
    std::vector<E> blk((size_t)1 << 22);
    for (ull e0 = 0; e0 < csr.m; e0 += blk.size()) {
        const ull cnt = std::min<ull>(blk.size(), csr.m - e0);
        parallel_for(cnt, [&](uint64_t b, uint64_t e, int) {
            for (uint64_t i = b; i < e; i++) blk[i] = csr.edgeAt(e0 + i);
        });
        write_exact(f, blk.data(), cnt * sizeof(E), path);
    }
}

template struct HostCsr<OutEdge>;
template struct HostCsr<OutEdgeWeighted>;
template std::shared_ptr<HostCsr<OutEdge>> read_csr_file<OutEdge>(const std::string &, bool);
template std::shared_ptr<HostCsr<OutEdgeWeighted>> read_csr_file<OutEdgeWeighted>(const std::string &, bool);
template void write_csr_file<OutEdge>(const HostCsr<OutEdge> &, const std::string &);
template void write_csr_file<OutEdgeWeighted>(const HostCsr<OutEdgeWeighted> &, const std::string &);

}  // namespace subway

// =================================================================== Graph
using subway::HostCsr;

// The host result array receives the run's final device-to-host copy of V
// values: pinned, that copy runs at the link's rate instead of through the
// runtime's staging buffers.  Ordinary memory if the pin is refused.
template <class T>
static T *alloc_result(uint n, bool &pinned)
{
    T *p = nullptr;
    if (hipHostMalloc((void **)&p, (size_t)(n ? n : 1) * sizeof(T), hipHostMallocPortable) == hipSuccess) {
        pinned = true;
        return p;
    }
    (void)hipGetLastError();
    pinned = false;
    return new T[n];
}

template <class T>
static void free_result(T *p, bool pinned)
{
    if (pinned)
        (void)hipHostFree(p);
    else
        delete[] p;
}

template <class E>
Graph<E>::Graph(std::string path, bool weighted)
    : graphFilePath(std::move(path)), isWeighted(weighted), isLarge(false), num_nodes(0), num_edges(0),
      nodePointer(nullptr), edgeList(nullptr), outDegree(nullptr), label1(nullptr), label2(nullptr),
      value(nullptr), d_outDegree(nullptr), d_value(nullptr), d_label1(nullptr), d_label2(nullptr),
      d_nodePointer(nullptr), d_edgeList(nullptr)
{
}

template <class E>
Graph<E>::Graph(std::shared_ptr<HostCsr<E>> c) : Graph<E>(std::string(), sizeof(E) == 8)
{
    csr = std::move(c);
    graphFormat = "memory";
    attach();
}

template <class E>
Graph<E>::~Graph()
{
    delete[] label1;
    delete[] label2;
    free_result(value, valuePinned);
    if (d_outDegree) (void)hipFree(d_outDegree);
    if (d_value) (void)hipFree(d_value);
    if (d_label1) (void)hipFree(d_label1);
    if (d_label2) (void)hipFree(d_label2);
    if (d_nodePointer) (void)hipFree(d_nodePointer);
}

template <class E>
std::string Graph<E>::GetFileExtension(std::string fileName)
{
    return subway::file_extension(fileName);
}

template <class E>
void Graph<E>::AssignW8(uint w8, uint index)
{
    if constexpr (sizeof(E) == 8) reinterpret_cast<uint *>(&edgeList[index])[1] = w8;
    (void)w8;
    (void)index;
}

template <class E>
void Graph<E>::ReadGraph()
{
    std::cout << "Reading the input graph from the following file:\n>> " << graphFilePath << std::endl;
    graphFormat = GetFileExtension(graphFilePath);
    csr = subway::read_csr_file<E>(graphFilePath);
    attach();
    std::cout << "Done reading.\n";
    std::cout << "Number of nodes = " << num_nodes << std::endl;
    std::cout << "Number of edges = " << num_edges << std::endl;
}

template <class E>
void Graph<E>::attach()
{
    if (!csr->pinned && !csr->splitDst)
        throw subway::Error(subway::ERR_ARG, "the graph's edge list is not in pinned memory (no GPU)");
    num_nodes = (uint)csr->n;
    num_edges = csr->m;
    isLarge = csr->large;
    nodePointer = csr->ptr;
    edgeList = csr->edges;
    outDegree = csr->degree;
    d_edgeList = csr->d_edges;
    label1 = new bool[num_nodes];
    label2 = new bool[num_nodes];
    value = alloc_result<uint>(num_nodes, valuePinned);
    SUBWAY_CHECK(hipMalloc(&d_outDegree, (size_t)num_nodes * sizeof(uint)));
    SUBWAY_CHECK(hipMalloc(&d_value, (size_t)num_nodes * sizeof(uint)));
    SUBWAY_CHECK(hipMalloc(&d_label1, (size_t)num_nodes * sizeof(bool)));
    SUBWAY_CHECK(hipMalloc(&d_label2, (size_t)num_nodes * sizeof(bool)));
    SUBWAY_CHECK(hipMalloc(&d_nodePointer, ((size_t)num_nodes + 1) * sizeof(ull)));
    SUBWAY_CHECK(hipMemcpy(d_nodePointer, nodePointer, ((size_t)num_nodes + 1) * sizeof(ull),
                           hipMemcpyHostToDevice));
    SUBWAY_CHECK(hipMemcpy(d_outDegree, outDegree, (size_t)num_nodes * sizeof(uint), hipMemcpyHostToDevice));
}

// ================================================================= GraphPR
template <class E>
GraphPR<E>::GraphPR(std::string path, bool weighted)
    : graphFilePath(std::move(path)), isWeighted(weighted), isLarge(false), num_nodes(0), num_edges(0),
      nodePointer(nullptr), edgeList(nullptr), outDegree(nullptr), value(nullptr), delta(nullptr),
      d_outDegree(nullptr), d_value(nullptr), d_delta(nullptr), d_nodePointer(nullptr), d_edgeList(nullptr)
{
}

template <class E>
GraphPR<E>::GraphPR(std::shared_ptr<HostCsr<E>> c) : GraphPR<E>(std::string(), sizeof(E) == 8)
{
    csr = std::move(c);
    graphFormat = "memory";
    attach();
}

template <class E>
GraphPR<E>::~GraphPR()
{
    free_result(value, valuePinned);
    delete[] delta;
    if (d_outDegree) (void)hipFree(d_outDegree);
    if (d_value) (void)hipFree(d_value);
    if (d_delta) (void)hipFree(d_delta);
    if (d_nodePointer) (void)hipFree(d_nodePointer);
}

template <class E>
std::string GraphPR<E>::GetFileExtension(std::string fileName)
{
    return subway::file_extension(fileName);
}

template <class E>
void GraphPR<E>::AssignW8(uint w8, uint index)
{
    if constexpr (sizeof(E) == 8) reinterpret_cast<uint *>(&edgeList[index])[1] = w8;
    (void)w8;
    (void)index;
}

template <class E>
void GraphPR<E>::ReadGraph()
{
    std::cout << "Reading the input graph from the following file:\n>> " << graphFilePath << std::endl;
    graphFormat = GetFileExtension(graphFilePath);
    csr = subway::read_csr_file<E>(graphFilePath);
    attach();
    std::cout << "Done reading.\n";
    std::cout << "Number of nodes = " << num_nodes << std::endl;
    std::cout << "Number of edges = " << num_edges << std::endl;
}

template <class E>
void GraphPR<E>::attach()
{
    if (!csr->pinned && !csr->splitDst)
        throw subway::Error(subway::ERR_ARG, "the graph's edge list is not in pinned memory (no GPU)");
    num_nodes = (uint)csr->n;
    num_edges = csr->m;
    isLarge = csr->large;
    nodePointer = csr->ptr;
    edgeList = csr->edges;
    outDegree = csr->degree;
    d_edgeList = csr->d_edges;
    value = alloc_result<float>(num_nodes, valuePinned);
    delta = new float[num_nodes];
    SUBWAY_CHECK(hipMalloc(&d_outDegree, (size_t)num_nodes * sizeof(uint)));
    SUBWAY_CHECK(hipMalloc(&d_value, (size_t)num_nodes * sizeof(float)));
    SUBWAY_CHECK(hipMalloc(&d_delta, (size_t)num_nodes * sizeof(float)));
    SUBWAY_CHECK(hipMalloc(&d_nodePointer, ((size_t)num_nodes + 1) * sizeof(ull)));
    SUBWAY_CHECK(hipMemcpy(d_nodePointer, nodePointer, ((size_t)num_nodes + 1) * sizeof(ull),
                           hipMemcpyHostToDevice));
    SUBWAY_CHECK(hipMemcpy(d_outDegree, outDegree, (size_t)num_nodes * sizeof(uint), hipMemcpyHostToDevice));
}

// shared/graph.cu:397-401
template class Graph<OutEdge>;
template class Graph<OutEdgeWeighted>;
template class GraphPR<OutEdge>;
template class GraphPR<OutEdgeWeighted>;
