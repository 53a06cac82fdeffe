/*
 * oracle.c -- serial CPU oracles for Subway's algorithms and subgraph
 * pipeline.  Test infrastructure only; see oracle.h for what pins each one.
 *
 * Update rules restated from the reference:
 *   BFS   cand = level[u] + 1, keep min        shared/gpu_kernels.cu:9-58
 *   CC    cand = label[u], keep min            shared/gpu_kernels.cu:60-108
 *   SSSP  cand = dist[u] + w, keep min         shared/gpu_kernels.cu:111-159
 *   SSWP  cand = min(width[u], w), keep max    shared/gpu_kernels.cu:161-209
 *   PR    delta-push, d/deg*0.85 per out-edge  shared/gpu_kernels.cu:211-254
 * Initial values follow the apps: BFS/SSSP INF with source 0
 * (subway/bfs-sync.cu:28-36, sssp-async.cu:38-44), SSWP 0 with source INF
 * (sswp-sync.cu:28-36), CC label = id (cc-sync.cu:29-34), PR value 0 and
 * delta 0.15 (pr-sync.cu:33-40).
 * The label algorithms reach a unique fixpoint, so they are computed here with
 * the textbook algorithm for that fixpoint (queue BFS, Dijkstra, widest-path
 * Dijkstra, min-ancestor reachability).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- BFS -- */
void oracle_bfs(uint64_t n, const uint64_t *ptr, const uint32_t *adj,
                uint32_t src, uint32_t *level)
{
    for (uint64_t i = 0; i < n; i++) level[i] = ORACLE_INF;
    if (src >= n) return;
    uint32_t *queue = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint64_t head = 0, tail = 0;
    level[src] = 0;
    queue[tail++] = src;
    while (head < tail) {
        uint32_t u = queue[head++];
        for (uint64_t e = ptr[u]; e < ptr[u + 1]; e++) {
            uint32_t v = adj[e];
            if (level[v] == ORACLE_INF) {
                level[v] = level[u] + 1;
                queue[tail++] = v;
            }
        }
    }
    free(queue);
}

/* ------------------------------------------------------ binary heap ---- */
typedef struct { uint64_t key; uint32_t v; } hitem;
typedef struct { hitem *a; uint64_t size, cap; int maxheap; } heap;

static int hless(const heap *h, uint64_t i, uint64_t j)
{
    return h->maxheap ? (h->a[i].key > h->a[j].key) : (h->a[i].key < h->a[j].key);
}
static void hpush(heap *h, uint64_t key, uint32_t v)
{
    if (h->size == h->cap) {
        h->cap = h->cap ? h->cap * 2 : 1024;
        h->a = (hitem *)realloc(h->a, sizeof(hitem) * h->cap);
    }
    uint64_t i = h->size++;
    h->a[i].key = key; h->a[i].v = v;
    while (i > 0) {
        uint64_t p = (i - 1) / 2;
        if (!hless(h, i, p)) break;
        hitem t = h->a[i]; h->a[i] = h->a[p]; h->a[p] = t;
        i = p;
    }
}
static hitem hpop(heap *h)
{
    hitem top = h->a[0];
    h->a[0] = h->a[--h->size];
    uint64_t i = 0;
    for (;;) {
        uint64_t l = 2 * i + 1, r = l + 1, b = i;
        if (l < h->size && hless(h, l, b)) b = l;
        if (r < h->size && hless(h, r, b)) b = r;
        if (b == i) break;
        hitem t = h->a[i]; h->a[i] = h->a[b]; h->a[b] = t;
        i = b;
    }
    return top;
}

/* --------------------------------------------------------------- SSSP -- */
void oracle_sssp(uint64_t n, const uint64_t *ptr, const uint32_t *adj,
                 const uint32_t *w, uint32_t src, uint32_t *dist)
{
    uint64_t *d = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    for (uint64_t i = 0; i < n; i++) d[i] = UINT64_MAX;
    heap h = {0};
    if (src < n) { d[src] = 0; hpush(&h, 0, src); }
    while (h.size) {
        hitem it = hpop(&h);
        if (it.key != d[it.v]) continue;
        uint32_t u = it.v;
        for (uint64_t e = ptr[u]; e < ptr[u + 1]; e++) {
            uint64_t c = d[u] + w[e];
            if (c < d[adj[e]]) { d[adj[e]] = c; hpush(&h, c, adj[e]); }
        }
    }
    for (uint64_t i = 0; i < n; i++)
        dist[i] = (d[i] >= ORACLE_INF) ? ORACLE_INF : (uint32_t)d[i];
    free(d); free(h.a);
}

/* --------------------------------------------------------------- SSWP -- */
void oracle_sswp(uint64_t n, const uint64_t *ptr, const uint32_t *adj,
                 const uint32_t *w, uint32_t src, uint32_t *width)
{
    for (uint64_t i = 0; i < n; i++) width[i] = 0;
    if (src >= n) return;
    heap h = {0};
    h.maxheap = 1;
    width[src] = ORACLE_INF;
    hpush(&h, ORACLE_INF, src);
    while (h.size) {
        hitem it = hpop(&h);
        if (it.key != width[it.v]) continue;
        uint32_t u = it.v;
        for (uint64_t e = ptr[u]; e < ptr[u + 1]; e++) {
            uint32_t c = width[u] < w[e] ? width[u] : w[e];
            if (c > width[adj[e]]) { width[adj[e]] = c; hpush(&h, c, adj[e]); }
        }
    }
    free(h.a);
}

/* ----------------------------------------------------------------- CC -- */
/* label[v] = min{u : u reaches v}; sweeping u upward, each BFS stops at
 * vertices already claimed by a smaller root (their reach is claimed too). */
void oracle_cc(uint64_t n, const uint64_t *ptr, const uint32_t *adj,
               uint32_t *label)
{
    const uint32_t UNSET = 0xFFFFFFFFu;
    for (uint64_t i = 0; i < n; i++) label[i] = UNSET;
    uint32_t *queue = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint64_t r = 0; r < n; r++) {
        if (label[r] != UNSET) continue;
        uint64_t head = 0, tail = 0;
        label[r] = (uint32_t)r;
        queue[tail++] = (uint32_t)r;
        while (head < tail) {
            uint32_t u = queue[head++];
            for (uint64_t e = ptr[u]; e < ptr[u + 1]; e++) {
                uint32_t v = adj[e];
                if (label[v] == UNSET) { label[v] = (uint32_t)r; queue[tail++] = v; }
            }
        }
    }
    free(queue);
}

/* ----------------------------------------------------------------- PR -- */
uint64_t oracle_pr(uint64_t n, const uint64_t *ptr, const uint32_t *adj,
                   double init, double damping, double acc, double *value,
                   double *delta)
{
    uint32_t *queue = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint8_t *inq = (uint8_t *)calloc(n ? n : 1, 1);
    uint64_t head = 0, count = 0, pushes = 0;
    for (uint64_t i = 0; i < n; i++) {
        value[i] = 0.0;
        delta[i] = init;
        if (init > acc) { queue[(head + count) % n] = (uint32_t)i; count++; inq[i] = 1; }
    }
    while (count) {
        uint32_t u = queue[head];
        head = (head + 1) % n; count--;
        inq[u] = 0;
        double d = delta[u];
        if (!(d > acc)) continue;
        pushes++;
        value[u] += d;
        delta[u] -= d;
        uint64_t deg = ptr[u + 1] - ptr[u];
        if (deg == 0) continue;
        double share = d / (double)deg * damping;
        for (uint64_t e = ptr[u]; e < ptr[u + 1]; e++) {
            uint32_t v = adj[e];
            delta[v] += share;
            if (!inq[v] && delta[v] > acc) {
                queue[(head + count) % n] = v; count++; inq[v] = 1;
            }
        }
    }
    free(queue); free(inq);
    return pushes;
}

/* ----------------------------------------------------- generator ------ */
/* shared/subgraph_generator.cu:10-60,128-162: flag, exclusive scans of flags
 * and masked degrees, scatter compaction in ascending id order; the pointer
 * array carries one extra entry holding the active edge total. */
uint64_t oracle_generate(uint64_t n, const uint8_t *active,
                         const uint32_t *degree, uint32_t *activeNodes,
                         uint64_t *activeNodesPointer)
{
    uint64_t a = 0, edges = 0;
    for (uint64_t v = 0; v < n; v++) {
        if (!active[v]) continue;
        activeNodes[a] = (uint32_t)v;
        activeNodesPointer[a] = edges;
        edges += degree[v];
        a++;
    }
    activeNodesPointer[a] = edges;
    return a;
}

/* ----------------------------------------------------- partitioner ---- */
/* shared/partitioner.cu:12-91: contiguous ranges of the active list, each no
 * larger than the budget.  Restated with an exact largest-fit search, and a
 * vertex larger than the budget gets a partition of its own instead of the
 * reference's endless loop. */
uint64_t oracle_partition(uint64_t numActive, const uint64_t *p,
                          uint64_t budget, uint64_t max_parts,
                          uint64_t *fromNode, uint64_t *nodeCount,
                          uint64_t *fromEdge, uint64_t *edgeCount)
{
    uint64_t from = 0, np = 0;
    while (from < numActive) {
        /* largest to with p[to] - p[from] <= budget */
        uint64_t lo = from, hi = numActive;
        while (lo < hi) {
            uint64_t mid = lo + (hi - lo + 1) / 2;
            if (p[mid] - p[from] <= budget) lo = mid; else hi = mid - 1;
        }
        uint64_t to = lo;
        if (to == from) to = from + 1; /* oversized vertex */
        if (np < max_parts) {
            fromNode[np] = from;
            nodeCount[np] = to - from;
            fromEdge[np] = p[from];
            edgeCount[np] = p[to] - p[from];
        }
        np++;
        from = to;
    }
    return np;
}
