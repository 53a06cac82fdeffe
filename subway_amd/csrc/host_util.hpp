// host_util.hpp -- small host-side helpers: a thread-count policy and a
// blocked parallel_for over std::thread.
#ifndef SUBWAY_HOST_UTIL_HPP
#define SUBWAY_HOST_UTIL_HPP

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <thread>
#include <vector>

namespace subway {

// Host worker threads: SUBWAY_HOST_THREADS, else OMP_NUM_THREADS, else the
// hardware concurrency; never more than 256.
inline int host_threads()
{
    const char *s = std::getenv("SUBWAY_HOST_THREADS");
    if (!s || !*s) s = std::getenv("OMP_NUM_THREADS");
    int n = s && *s ? std::atoi(s) : 0;
    if (n <= 0) n = (int)std::thread::hardware_concurrency();
    if (n <= 0) n = 1;
    return std::min(n, 256);
}

// fn(begin, end, thread_index) over [0, n) split into contiguous blocks.
template <class F>
void parallel_for(uint64_t n, F fn, int threads = 0)
{
    if (threads <= 0) threads = host_threads();
    if (n < 4096 || threads == 1) {
        fn((uint64_t)0, n, 0);
        return;
    }
    uint64_t t = std::min<uint64_t>((uint64_t)threads, n);
    std::vector<std::thread> pool;
    pool.reserve(t);
    for (uint64_t i = 0; i < t; i++) {
        uint64_t b = n * i / t, e = n * (i + 1) / t;
        pool.emplace_back([=, &fn] { fn(b, e, (int)i); });
    }
    for (auto &th : pool) th.join();
}

}  // namespace subway

#endif  // SUBWAY_HOST_UTIL_HPP
