/*
 * omp_baseline.c -- OpenMP build of Subway's update rules over the same CSR,
 * used only as the CPU baseline that is timed beside the GPU path and as a
 * second cross-check in the tests.  It is not the product.
 *
 * Same initial values and relaxations as the GPU kernels
 * (shared/gpu_kernels.cu:9-254 for the rules, the subway/ *-sync.cu apps for the
 * initial values): CAS-based uint32 min/max for BFS/CC/SSSP/SSWP and an FP32
 * atomic add for PageRank pushes of (float)((d/deg)*0.85).  There are no
 * host-device transfers; every round is a bulk-synchronous frontier sweep in
 * which values written earlier in the same round are already visible, so it
 * reaches the same fixpoint as the sync and async GPU schedules.
 */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define INF 4294967294u

enum { ALGO_BFS = 0, ALGO_CC = 1, ALGO_SSSP = 2, ALGO_SSWP = 3 };

typedef struct { uint32_t *v; uint64_t n, cap; } vec;

static void vpush(vec *q, uint32_t x)
{
    if (q->n == q->cap) {
        q->cap = q->cap ? q->cap * 2 : 4096;
        q->v = (uint32_t *)realloc(q->v, sizeof(uint32_t) * q->cap);
    }
    q->v[q->n++] = x;
}

/* Concatenate per-thread lists into `out`; returns the new length. */
static uint64_t gather_lists(vec *lists, int nt, uint32_t *out)
{
    uint64_t total = 0;
    for (int t = 0; t < nt; t++) {
        memcpy(out + total, lists[t].v, sizeof(uint32_t) * lists[t].n);
        total += lists[t].n;
        lists[t].n = 0;
    }
    return total;
}

/* Weights come from w (uint32) or, for graphs held as split arrays, w8
 * (one byte per edge); exactly one of them is used. */
static inline uint32_t weight_of(const uint32_t *w, const uint8_t *w8, uint64_t e)
{
    return w ? w[e] : (uint32_t)w8[e];
}

static uint64_t label_run(int algo, uint64_t n, const uint64_t *ptr, const uint32_t *adj, const uint32_t *w,
                          const uint8_t *w8, uint32_t src, uint32_t *value, uint64_t *edges_out);

/* Returns the number of rounds; *edges_out counts relaxed edges. */
uint64_t omp_label_run(int algo, uint64_t n, const uint64_t *ptr,
                       const uint32_t *adj, const uint32_t *w, uint32_t src,
                       uint32_t *value, uint64_t *edges_out)
{
    return label_run(algo, n, ptr, adj, w, NULL, src, value, edges_out);
}

/* The same over byte weights (a graph held as split destination and weight
 * arrays, e.g. the 8.59 G-edge R-MAT). */
uint64_t omp_label_run_w8(int algo, uint64_t n, const uint64_t *ptr,
                          const uint32_t *adj, const uint8_t *w8, uint32_t src,
                          uint32_t *value, uint64_t *edges_out)
{
    return label_run(algo, n, ptr, adj, NULL, w8, src, value, edges_out);
}

static uint64_t label_run(int algo, uint64_t n, const uint64_t *ptr, const uint32_t *adj, const uint32_t *w,
                          const uint8_t *w8, uint32_t src, uint32_t *value, uint64_t *edges_out)
{
    uint32_t *front = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint8_t *inNext = (uint8_t *)calloc(n ? n : 1, 1);
    uint64_t fsize = 0, rounds = 0, edges = 0;
    int nt = omp_get_max_threads();
    vec *lists = (vec *)calloc(nt, sizeof(vec));

    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) {
        if (algo == ALGO_CC) value[i] = (uint32_t)i;
        else if (algo == ALGO_SSWP) value[i] = 0;
        else value[i] = INF;
    }
    if (algo == ALGO_CC) {
        for (uint64_t i = 0; i < n; i++) front[i] = (uint32_t)i;
        fsize = n;
    } else if (src < n) {
        value[src] = (algo == ALGO_SSWP) ? INF : 0;
        front[0] = src;
        fsize = 1;
    }

    while (fsize) {
        rounds++;
        #pragma omp parallel reduction(+:edges)
        {
            vec *mine = &lists[omp_get_thread_num()];
            #pragma omp for schedule(dynamic, 64)
            for (uint64_t i = 0; i < fsize; i++) {
                uint32_t u = front[i];
                uint32_t val = __atomic_load_n(&value[u], __ATOMIC_RELAXED);
                for (uint64_t e = ptr[u]; e < ptr[u + 1]; e++) {
                    uint32_t v = adj[e];
                    uint32_t cand;
                    int better;
                    uint32_t old = __atomic_load_n(&value[v], __ATOMIC_RELAXED);
                    if (algo == ALGO_BFS) cand = val + 1;
                    else if (algo == ALGO_CC) cand = val;
                    else if (algo == ALGO_SSSP) cand = val + weight_of(w, w8, e);
                    else {
                        const uint32_t we = weight_of(w, w8, e);
                        cand = val < we ? val : we;
                    }
                    edges++;
                    for (;;) {
                        better = (algo == ALGO_SSWP) ? (cand > old) : (cand < old);
                        if (!better) break;
                        if (__atomic_compare_exchange_n(&value[v], &old, cand, 0,
                                                        __ATOMIC_RELAXED, __ATOMIC_RELAXED))
                            break;
                    }
                    if (better && !__atomic_exchange_n(&inNext[v], 1, __ATOMIC_RELAXED))
                        vpush(mine, v);
                }
            }
        }
        fsize = gather_lists(lists, nt, front);
        #pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < fsize; i++) inNext[front[i]] = 0;
    }
    for (int t = 0; t < nt; t++) free(lists[t].v);
    free(lists); free(front); free(inNext);
    if (edges_out) *edges_out = edges;
    return rounds;
}

static void atomic_add_f32(float *addr, float x)
{
    uint32_t *p = (uint32_t *)addr;
    uint32_t old = __atomic_load_n(p, __ATOMIC_RELAXED), nw;
    for (;;) {
        float f;
        memcpy(&f, &old, 4);
        f += x;
        memcpy(&nw, &f, 4);
        if (__atomic_compare_exchange_n(p, &old, nw, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
            return;
    }
}

/* Delta-push PageRank with FP32 state, pushes of (float)((d/deg)*0.85)
 * (shared/gpu_kernels.cu:459-502).  Returns the number of rounds. */
uint64_t omp_pr_run(uint64_t n, const uint64_t *ptr, const uint32_t *adj,
                    float init, float acc, float *value, float *delta,
                    uint64_t *edges_out)
{
    uint32_t *front = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint8_t *inNext = (uint8_t *)calloc(n ? n : 1, 1);
    uint64_t fsize = 0, rounds = 0, edges = 0;
    int nt = omp_get_max_threads();
    vec *lists = (vec *)calloc(nt, sizeof(vec));
    for (uint64_t i = 0; i < n; i++) {
        value[i] = 0.0f;
        delta[i] = init;
        if (init > acc) front[fsize++] = (uint32_t)i;
    }
    while (fsize) {
        rounds++;
        #pragma omp parallel reduction(+:edges)
        {
            vec *mine = &lists[omp_get_thread_num()];
            #pragma omp for schedule(dynamic, 64)
            for (uint64_t i = 0; i < fsize; i++) {
                uint32_t u = front[i];
                float d;
                uint32_t bits = __atomic_load_n((uint32_t *)&delta[u], __ATOMIC_RELAXED);
                memcpy(&d, &bits, 4);
                if (!(d > acc)) continue;
                value[u] += d;
                uint64_t deg = ptr[u + 1] - ptr[u];
                if (deg) {
                    float share = (float)((double)(d / (float)deg) * 0.85);
                    for (uint64_t e = ptr[u]; e < ptr[u + 1]; e++) {
                        uint32_t v = adj[e];
                        atomic_add_f32(&delta[v], share);
                        edges++;
                        float dv;
                        uint32_t b2 = __atomic_load_n((uint32_t *)&delta[v], __ATOMIC_RELAXED);
                        memcpy(&dv, &b2, 4);
                        if (dv > acc && !__atomic_exchange_n(&inNext[v], 1, __ATOMIC_RELAXED))
                            vpush(mine, v);
                    }
                }
                atomic_add_f32(&delta[u], -d);
                float du;
                uint32_t b3 = __atomic_load_n((uint32_t *)&delta[u], __ATOMIC_RELAXED);
                memcpy(&du, &b3, 4);
                if (du > acc && !__atomic_exchange_n(&inNext[u], 1, __ATOMIC_RELAXED))
                    vpush(mine, u);
            }
        }
        fsize = gather_lists(lists, nt, front);
        #pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < fsize; i++) inNext[front[i]] = 0;
    }
    for (int t = 0; t < nt; t++) free(lists[t].v);
    free(lists); free(front); free(inNext);
    if (edges_out) *edges_out = edges;
    return rounds;
}

int omp_num_threads(void) { return omp_get_max_threads(); }
