// rmat.cpp -- reproducible R-MAT (Kronecker, no noise, no vertex
// permutation) graph generator writing CSR straight into pinned host memory.
//
// Edge e (0 <= e < edgeFactor * 2^scale) draws one uniform per level from a
// counter-based hash of (seed, e, level), so the output does not depend on
// the thread count.  Level l picks a quadrant with probabilities (a, b, c, d)
// and sets bit (scale-1-l) of the source / destination.  Self-loops are
// dropped, duplicates kept.  Weights are uniform on [wmin, wmax] from an
// independent hash stream.  With `symmetrize` every kept edge (u,v,w) is
// stored as both (u,v,w) and (v,u,w).  The CSR is built by a stable counting
// sort by source (edge order inside a vertex = generation order), done as a
// bucket pass over source high bits and a per-bucket counting sort.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "host_util.hpp"
#include "subway/globals.hpp"
#include "subway/gpu_error_check.hpp"
#include "subway/graph.hpp"

namespace subway {

static inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct RmatParams {
    int scale;
    uint64_t nominal;     // edgeFactor * 2^scale
    uint32_t ta, tab, tabc;  // cumulative thresholds on 2^32
    uint64_t seed;
    uint32_t wmin, wspan;
};

// Returns false for a self-loop.
static inline bool rmat_edge(const RmatParams &p, uint64_t e, uint32_t &src, uint32_t &dst)
{
    uint32_t s = 0, d = 0;
    const uint64_t key = mix64(p.seed ^ 0x5DEECE66Dull);
    for (int l = 0; l < p.scale; l += 2) {
        uint64_t h = mix64(key ^ mix64((e << 5) | (uint64_t)(l >> 1)));
        for (int k = 0; k < 2 && l + k < p.scale; k++) {
            uint32_t r = (uint32_t)(h >> (32 * k));
            uint32_t bit = 1u << (p.scale - 1 - (l + k));
            if (r < p.ta) {
            } else if (r < p.tab) {
                d |= bit;
            } else if (r < p.tabc) {
                s |= bit;
            } else {
                s |= bit;
                d |= bit;
            }
        }
    }
    src = s;
    dst = d;
    return s != d;
}

static inline uint32_t rmat_weight(const RmatParams &p, uint64_t e)
{
    uint64_t h = mix64(mix64(p.seed ^ 0xC2B2AE3D27D4EB4Full) ^ e);
    return p.wmin + (uint32_t)(h % p.wspan);
}

// Graphs of 2^31 nominal edges and more take minutes to build: each pass
// reports its end on stderr.
struct RmatLog {
    bool on;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    void pass(const char *what) const
    {
        if (!on) return;
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::fprintf(stderr, "[rmat] %s done at %.1f s\n", what, s);
        std::fflush(stderr);
    }
};

template <class E>
std::shared_ptr<HostCsr<E>> rmat_csr(int scale, int edgeFactor, double a, double b, double c, uint64_t seed,
                                     bool symmetrize, uint32_t wmin, uint32_t wmax, bool splitLayout)
{
    if (scale < 1 || scale > 31) throw Error(ERR_ARG, "R-MAT scale must be in [1, 31]");
    if (edgeFactor < 1) throw Error(ERR_ARG, "R-MAT edge factor must be positive");
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) throw Error(ERR_ARG, "R-MAT probabilities invalid");
    if (wmin < 1 || wmax < wmin) throw Error(ERR_ARG, "R-MAT weight range invalid");
    RmatParams p;
    p.scale = scale;
    p.nominal = (uint64_t)edgeFactor << scale;
    auto thr = [](double x) -> uint32_t {
        double v = x * 4294967296.0;
        return v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)v;
    };
    p.ta = thr(a);
    p.tab = thr(a + b);
    p.tabc = thr(a + b + c);
    p.seed = seed;
    p.wmin = wmin;
    p.wspan = wmax - wmin + 1;
    const uint64_t n = 1ull << scale;
    const int T = host_threads();
    // buckets by the top bits of the source id
    const int bbits = std::min(scale, 10);
    const uint64_t NB = 1ull << bbits;
    const int bshift = scale - bbits;
    const int copies = symmetrize ? 2 : 1;
    const RmatLog log{p.nominal >= (1ull << 31)};

    // pass 1: per-thread bucket counts
    std::vector<uint64_t> cnt((size_t)T * NB, 0);
    std::vector<uint64_t> kept(T, 0);
    parallel_for(p.nominal, [&](uint64_t b0, uint64_t b1, int t) {
        uint64_t *c = &cnt[(size_t)t * NB];
        uint64_t k = 0;
        for (uint64_t e = b0; e < b1; e++) {
            uint32_t s, d;
            if (!rmat_edge(p, e, s, d)) continue;
            c[s >> bshift]++;
            if (symmetrize) c[d >> bshift]++;
            k++;
        }
        kept[t] = k;
    }, T);
    log.pass("pass 1 of 3 (bucket counts)");
    // bucket offsets: bucket-major, then thread order (keeps generation order)
    std::vector<uint64_t> off((size_t)T * NB);
    std::vector<uint64_t> bstart(NB + 1);
    uint64_t run = 0;
    for (uint64_t bk = 0; bk < NB; bk++) {
        bstart[bk] = run;
        for (int t = 0; t < T; t++) {
            off[(size_t)t * NB + bk] = run;
            run += cnt[(size_t)t * NB + bk];
        }
    }
    bstart[NB] = run;
    const uint64_t m = run;
    (void)copies;

    struct Rec { uint32_t s, d, w; };
    std::unique_ptr<Rec[]> tmp(new Rec[m ? m : 1]);
    // pass 2: regenerate and scatter into buckets
    parallel_for(p.nominal, [&](uint64_t b0, uint64_t b1, int t) {
        uint64_t *o = &off[(size_t)t * NB];
        for (uint64_t e = b0; e < b1; e++) {
            uint32_t s, d;
            if (!rmat_edge(p, e, s, d)) continue;
            uint32_t w = rmat_weight(p, e);
            tmp[o[s >> bshift]++] = Rec{s, d, w};
            if (symmetrize) tmp[o[d >> bshift]++] = Rec{d, s, w};
        }
    }, T);
    log.pass("pass 2 of 3 (edges into buckets)");

    auto csr = std::make_shared<HostCsr<E>>();
    // A large weighted graph whose ids and weight span do not fit one word
    // but whose weights fit a byte moves in the split format only: its
    // pinned memory is the split arrays (5 bytes per edge), and the 8-byte
    // list stays in ordinary memory for host-side use.  Pinning both does
    // not fit a host that pins about 70 GB per process (scale 28: 68.7 GB
    // of 8-byte edges).
    bool split = false;
    if constexpr (sizeof(E) == 8) {
        const bool packs = (uint64_t)(wmax - wmin) < (1ull << (32 - std::min(scale, 31)));
        split = !packs && wmax <= 0xFFu && (uint64_t)m * sizeof(E) >= (1ull << 34);
        if (splitLayout) {
            if (wmax > 0xFFu) throw Error(ERR_ARG, "the split layout needs weights that fit a byte");
            split = true;
        }
    }
    csr->allocate(n, m, !split);
    if (split && !csr->allocateSplit())
        throw Error(ERR_NOMEM, "pinned allocation of the split edge arrays failed");
    csr->large = m > 0xFFFFFFFFull;
    // pass 3: per bucket, stable counting sort by source
    std::atomic<uint64_t> nextBucket{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < T; t++) {
        pool.emplace_back([&] {
            std::vector<uint64_t> deg((size_t)1 << bshift);
            for (;;) {
                uint64_t bk = nextBucket.fetch_add(1);
                if (bk >= NB) break;
                const uint64_t v0 = bk << bshift, nv = 1ull << bshift;
                std::fill(deg.begin(), deg.end(), 0);
                for (uint64_t i = bstart[bk]; i < bstart[bk + 1]; i++) deg[tmp[i].s - v0]++;
                uint64_t r = bstart[bk];
                for (uint64_t v = 0; v < nv; v++) {
                    csr->ptr[v0 + v] = r;
                    uint64_t dg = deg[v];
                    deg[v] = r;
                    r += dg;
                }
                for (uint64_t i = bstart[bk]; i < bstart[bk + 1]; i++) {
                    const Rec &rc = tmp[i];
                    uint64_t at = deg[rc.s - v0]++;
                    csr->edges[at].end = rc.d;
                    if constexpr (sizeof(E) == 8) reinterpret_cast<uint32_t *>(&csr->edges[at])[1] = rc.w;
                    if (split) {
                        csr->splitDst[at] = rc.d;
                        csr->splitW[at] = (uint8_t)rc.w;
                    }
                }
            }
        });
    }
    for (auto &th : pool) th.join();
    log.pass("pass 3 of 3 (CSR)");
    tmp.reset();
    csr->finalize();
    return csr;
}

template std::shared_ptr<HostCsr<OutEdge>> rmat_csr<OutEdge>(int, int, double, double, double, uint64_t, bool,
                                                             uint32_t, uint32_t, bool);
template std::shared_ptr<HostCsr<OutEdgeWeighted>> rmat_csr<OutEdgeWeighted>(int, int, double, double, double,
                                                                             uint64_t, bool, uint32_t, uint32_t,
                                                                             bool);

}  // namespace subway
