"""Benchmark of the out-of-GPU-memory pipeline on one MI355X.

Default workload (the configuration the headline metric is quoted on):
SSSP-async on a weighted R-MAT graph, scale 28, edge factor 32
(268,435,456 vertices, ~8.59 G edges, a=0.57 b=0.19 c=0.19, weights
U[1,64], self-loops dropped), from the highest out-degree vertex, with the
device edge budget capped at 16 GiB, under the learned scheduler and under
the plain Subway policy.  The graph is built on the GPU (the host
generator's CSR, byte for byte) into pinned host memory, where every run
reads it: the path's job is to move it over the host link.  By default the
graph is then renumbered on the GPU by descending in-degree (--order
natural keeps the generator's ids), a one-time step reported as
config.reorder_s; sources and values stay in the original numbering.
Then (--blocks 1, the default) every row is sorted by destination and the
block-compressed copy of the rows is built on the GPU (config.blocks_s, its
bytes per edge in config.block_bytes_per_edge): whole id ranges then move
in that layout; every policy variant uses it.

A step is one full run from the initial state to the fixpoint (first
generate .. values in the session's host result array).  value = input
edges / mean step time of the learned policy (GTEPS); wall clock over the
timed steps, all variants interleaved step by step.  Variants (an ablation
of the learned schedule): "plain" (the reference's schedule), "rule" (the
plain schedule plus in-place visits of id ranges a buffer still holds, no
network), "network" (the network's decisions without in-place visits) and
"learned" (network plus in-place visits); their values must be identical.
Each timed step records HIP events around every rounds-kernel launch (on
its stream) and every transfer, from which the roofline entry and the
achieved H2D rate are computed.  The line also carries the network's
next-volume prediction error and the transfer lower bound every correct
schedule of this workload obeys.

CPU baseline: the OpenMP build of the same update rules on the same CSR
(the pinned split arrays, no copy), OMP_PROC_BIND=close with the process's
OMP_NUM_THREADS, median of 3 runs after one warm-up, values checked against
the GPU's.

--config c2|c3cc|c3pr|c4 runs the other configurations (c4 = 64-source
batched SSSP against 64 sequential runs).  Prints one JSON line on stdout
and exits non-zero when any comparison fails; progress goes to stderr.
"""
import argparse
import json
import os
import platform
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the OpenMP runtime reads these when the baseline library loads
os.environ.setdefault("OMP_PROC_BIND", "close")
# the CPUs this process may run on, read before any OpenMP region binds the
# main thread to one of them
AFFINITY_CPUS = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()

HBM_PEAK_GBS = 8000.0
METRIC = "end-to-end GTEPS + H2D GB moved; learned vs plain policy, same exact result"
GiB = 1 << 30

CONFIGS = {
    "c2": dict(algo="sssp", scale=24, edge_factor=16, budget="1G", weighted=True, symmetrize=False),
    "c3cc": dict(algo="cc", scale=25, edge_factor=16, budget="2G", weighted=False, symmetrize=True),
    "c3pr": dict(algo="pr", scale=25, edge_factor=16, budget="2G", weighted=False, symmetrize=False),
    "c4": dict(algo="sssp", scale=26, edge_factor=16, budget="4G", weighted=True, symmetrize=False, batch=64),
    "c5": dict(algo="sssp", scale=28, edge_factor=32, budget="16G", weighted=True, symmetrize=False),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse_size(s):
    mul = {"K": 1 << 10, "M": 1 << 20, "G": 1 << 30}
    s = str(s)
    return int(float(s[:-1]) * mul[s[-1].upper()]) if s[-1].upper() in mul else int(s)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def profiled_traffic(kernel, workload_key):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summaries (FETCH_SIZE and WRITE_SIZE passes, each scaled by the
    copy-kernel calibration they record), or None."""
    pdir = os.path.join(ROOT, "profiles")
    try:
        names = sorted(os.listdir(pdir))
    except OSError:
        return None
    for name in names:
        if not name.endswith("hbm_counters.txt"):
            continue
        with open(os.path.join(pdir, name)) as f:
            for line in f:
                parts = line.split()
                if parts[:2] == ["traffic_bytes_per_launch", kernel] and " ".join(parts[2:-1]) == workload_key:
                    return int(parts[-1])
    return None


def workload_name(args, budget):
    kind = "weighted " if args.weighted else ""
    sym = "symmetrised " if args.symmetrize else ""
    src = ("" if args.algo in ("cc", "pr") else
           ", source = max out-degree vertex")
    thr = ", threshold 1e-6" if args.algo == "pr" else ""
    return (f"{args.algo}-{args.mode} on {sym}{kind}R-MAT scale {args.scale} ef {args.edge_factor}{src}{thr}, "
            f"device edge budget {budget} B ({budget / GiB:g} GiB)")


VARIANTS = {
    "plain": dict(policy="plain"),
    "rule": dict(policy="learned", theta=2.0, params=None),
    "network": dict(policy="learned", resident=False),
    "learned": dict(policy="learned"),
}


def measure(sw, np, g, args, src, kw, names, params):
    """Warm-up and timed steps of each variant, interleaved step by step so
    that all see the same machine state; each run is timed on its own.
    Returns {name: (summary, stats, values of the last step)}."""
    sessions = {}
    for name in names:
        t0 = time.perf_counter()
        v = dict(VARIANTS[name])
        if "params" not in v:
            v["params"] = params if v["policy"] == "learned" else None
        sessions[name] = sw.Session(g, args.algo, **kw, **v)
        log(f"{name}: session setup {1e3 * (time.perf_counter() - t0):.1f} ms")
    for i in range(args.warmup):
        for name in names:
            _, st = sessions[name].run(src, copy=False)
            log(f"{name} warmup {i}: {st.ms_total:.1f} ms, {st.global_iterations} global iterations")
    stats = {n: [] for n in names}
    wall = {n: 0.0 for n in names}
    vals = {}
    for i in range(args.steps):
        for name in names:
            t0 = time.perf_counter()
            v, st = sessions[name].run(src, copy=False)
            wall[name] += time.perf_counter() - t0
            stats[name].append(st)
            if i == args.steps - 1:
                vals[name] = np.array(v, copy=True)
            log(f"{name} step {i}: {st.ms_total:.1f} ms e2e, {st.global_iterations} iters, {st.partitions} partitions, "
                f"H2D {st.h2d_bytes / 1e9:.3f} GB, rounds {st.ms_rounds:.1f} ms ({st.rounds_launches} launches), "
                f"transfer {st.ms_transfer:.1f} ms, in-place visits {st.resident_visits}, "
                f"early read-back {st.early_readback} ({st.values_patched} values patched)")
    out = {}
    for name in names:
        ss = stats[name]
        for r in ss[-1].trace:
            log(f"  {name} iter {r['iteration']}: active {r['active_nodes']} v / {r['active_edges']} e, "
                f"{r['partitions']} partitions, H2D {r['h2d_bytes'] / 1e6:.1f} MB, {r['ms']:.2f} ms, "
                f"theta {r['theta']:.2f}, cap {r['rounds_cap']}{' (plain)' if r['on_plain'] else ''}")
        ms_step = wall[name] / args.steps * 1e3
        h2d = float(np.mean([s.h2d_bytes for s in ss]))
        xfer_ms = float(np.mean([s.ms_transfer for s in ss]))
        out[name] = ({
            "gteps": round(g.m / (ms_step / 1e3) / 1e9, 4),
            "ms_per_step": round(ms_step, 3),
            "e2e_ms_library": round(float(np.mean([s.ms_total for s in ss])), 3),
            "global_iterations": ss[-1].global_iterations,
            "partitions": ss[-1].partitions,
            "inner_rounds": ss[-1].inner_rounds,
            "in_place_visits": ss[-1].resident_visits,
            "h2d_gb_per_step": round(h2d / 1e9, 4),
            "h2d_blocked_gb_per_step": round(float(np.mean([s.h2d_blocked_bytes for s in ss])) / 1e9, 4),
            "blocked_units": ss[-1].blocked_units,
            "h2d_mb_per_iteration": [round(r["h2d_bytes"] / 1e6, 2) for r in ss[-1].trace],
            "h2d_gbps_achieved": round(h2d / (xfer_ms * 1e-3) / 1e9, 2) if xfer_ms > 0 else None,
            "ms_breakdown": {"rounds": round(float(np.mean([s.ms_rounds for s in ss])), 3),
                             "transfer": round(xfer_ms, 3),
                             "generate": round(float(np.mean([s.ms_generate for s in ss])), 3)},
            "mispredictions": ss[-1].mispredictions,
            "fallback_iteration": ss[-1].fallback_iteration,
            # 1: results copied to the host during the last iterations, then
            # the values that changed patched; 2: ... then copied in full
            "early_readback": ss[-1].early_readback,
            "values_patched": ss[-1].values_patched,
        }, ss, vals[name])
    for s in sessions.values():
        s.close()
    return out


def prediction_error(trace):
    """Head 0 predicts log2 of the next iteration's active-edge fraction
    (/ 32); the next iteration's feature 2 is the fraction it saw.  Returns
    per-iteration |log2(predicted / actual)| and the judge's count of
    predictions more than 2x off (iterations under 1e-4 of the edges,
    predicted and actual, are not judged)."""
    import math
    errs, judged, missed = [], 0, 0
    for a, b in zip(trace, trace[1:]):
        if a["on_plain"]:
            continue
        pred = 2.0 ** (32.0 * min(max(a["heads"][0], -1.0), 0.0))
        act = b["features"][2]
        errs.append(round(abs(math.log2((pred + 1e-6) / (act + 1e-6))), 3))
        if pred >= 1e-4 or act >= 1e-4:
            judged += 1
            missed += int(not 0.5 <= (pred + 1e-6) / (act + 1e-6) <= 2.0)
    return {"abs_log2_error_per_iteration": errs, "judged": judged, "missed_by_more_than_2x": missed}


def roofline(stats, args):
    """Roofline of the dominant kernel.  Persistent async rounds: one
    rounds_kernel launch per partition visit does every inner round (frontier
    + advance).  Algorithmic bytes, summed over the rounds: label algorithms
    22 B per labelled active vertex (id 4 + offset 8 + degree 4 + label r/w 2
    + value 4), s(E) + 4 B per relaxed edge (edge as buffered + destination
    value), 9 B per successful update (atomic RMW 8 + label 1); PageRank
    (FP64 residual) 48 B per pushing vertex (id 4 + offset 8 + degree 4 +
    residual read and clear 16 + running-sum RMW 16) and s(E) + 16 B per
    edge (FP64 atomic add RMW).  The frontier's scan of unlabelled items is
    reported apart (items_scanned, at the bytes each scanned item reads) and
    not counted.  Three kinds of launches are told apart, each priced with
    its own time and work: persistent launches over units in the block
    layout (rounds_kernel<BlockEdge>, s(E) = the blocks' bytes per edge),
    persistent launches over plain units (rounds_kernel, s(E) = the plain
    edge size), and stand-alone frontier + advance launch pairs; the kind
    with the most kernel time is reported."""
    S = lambda f: sum(getattr(s, f) for s in stats)
    eb = stats[-1].edge_bytes
    blk_e = S("blocked_edges")
    eb_blk = S("h2d_blocked_bytes") / blk_e if blk_e else eb
    blk_ms, blk_n = S("ms_rounds_blocked"), S("rounds_launches_blocked")
    pair_w, pair_n_ent, pair_upd = S("pair_work_edges"), S("pair_work_entries"), S("pair_updates")
    blk_w, blk_ent, blk_upd = S("blk_work_edges"), S("blk_work_entries"), S("blk_updates")
    kinds = {
        "rounds_kernel<BlockEdge>": (blk_ms, blk_n, blk_w, blk_ent, blk_upd, eb_blk),
        "rounds_kernel": (S("ms_rounds") - blk_ms, S("rounds_launches") - blk_n,
                          S("work_edges") - pair_w - blk_w, S("work_entries") - pair_n_ent - blk_ent,
                          S("updates") - pair_upd - blk_upd, eb),
        "advance_kernel+frontier_kernel": (S("ms_advance") + S("ms_frontier"), S("advance_launches"), pair_w,
                                           pair_n_ent, pair_upd, eb),
    }
    kern = max(kinds, key=lambda k: kinds[k][0] if kinds[k][1] else -1.0)
    k_ms, k_launches, work, ents, upd, seb = kinds[kern]
    if args.algo == "pr":
        alg_bytes = 48 * ents + (seb + 16) * work
    else:
        alg_bytes = 22 * ents + (seb + 4) * work + 9 * upd
    items = S("work_items")
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    key = (f"{args.algo}-{args.mode} rmat{args.scale}ef{args.edge_factor} budget{args.budget} learned"
           + (" blocks" if args.blocks else ""))
    traffic = profiled_traffic(kern, key)
    per_launch = alg_bytes / max(k_launches, 1)
    return {
        "bound": "hbm",
        "kernel": kern,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "traffic_over_algorithmic": round(traffic / per_launch, 2) if traffic and alg_bytes else None,
        "launches": k_launches,
        "avg_launch_us": round(k_ms / max(k_launches, 1) * 1e3, 2),
        "algorithmic_bytes_per_launch": int(per_launch),
        "items_scanned_per_launch": int(items / max(S("rounds_launches"), 1)),
        "edge_bytes": round(seb, 4),
        "profile_key": key,
        "by_kind_ms": {k: round(v[0], 3) for k, v in kinds.items()},
    }


def original_degrees(g):
    """Out-degrees indexed by original vertex id (a renumbered graph's
    arrays are in its internal numbering)."""
    deg = g.out_degree()
    o = g.order()
    return deg[o[0]] if o is not None else deg


def to_original(np, g, algo, vals):
    """Values of a run on a renumbered graph's own CSR (internal ids) in the
    original numbering; CC labels (smallest internal id of each component)
    become the smallest original id, as the library reports them."""
    o = g.order()
    if o is None:
        return vals
    perm, iperm = o
    out = vals[perm]
    if algo == "cc":
        # component key = internal label; canonical label = smallest original id with that key
        low = np.full(g.n, np.iinfo(np.uint32).max, np.uint32)
        np.minimum.at(low, out, np.arange(g.n, dtype=np.uint32))
        out = low[out]
    return out


def cpu_baseline(np, g, args, src, gpu_vals):
    """OpenMP build of the same rules on the same CSR (the graph's own host
    arrays, so a renumbered graph's internal numbering, its results mapped
    back), timed as one warm-up then the median of 3."""
    from tests import _oracle as O
    o = g.order()
    if o is not None:
        src = int(o[0][src])
    if args.weighted and not args.symmetrize:
        try:
            ptr, adj, w = g.split_view()
        except Exception:
            ptr, adj, w = g.csr()
    else:
        ptr, adj, w = g.csr()
    times, vals = [], None
    for i in range(args.cpu_runs + 1):
        t = time.perf_counter()
        if args.algo == "pr":
            vals, _, _ = O.omp_pr(ptr, adj, 1e-6)
        else:
            vals, _, _ = O.omp_label(args.algo, ptr, adj, w if args.algo in ("sssp", "sswp") else None, src)
        dt = time.perf_counter() - t
        log(f"cpu baseline {'warm-up' if i == 0 else 'run ' + str(i)}: {dt:.2f} s")
        if i:
            times.append(dt)
    med = statistics.median(times)
    vals = to_original(np, g, args.algo, vals)
    if args.algo == "pr":
        a = gpu_vals.astype(np.float64)
        same = bool(np.max(np.abs(a - vals) / np.maximum(np.abs(a), 1e-30)) < 1e-3)
    else:
        same = bool(np.array_equal(vals, gpu_vals))
    aff = AFFINITY_CPUS
    return {
        "value": round(g.m / med / 1e9, 4), "unit": "GTEPS", "cores": O.omp_threads(), "kind": "port",
        "sample": f"the full workload: one {args.algo} run of the OpenMP build on the same {g.m}-edge CSR and source, "
                  f"median of {len(times)} after one warm-up, {med:.2f} s",
        "seconds": [round(t, 3) for t in times],
        "nproc": os.cpu_count(), "affinity_cpus": aff, "cpu_model": cpu_model(),
        "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
        "same_result_as_gpu": same,
    }


def batch_bench(sw, np, g, args, budget, peak):
    """K-source batched SSSP against K sequential single-source runs."""
    k = args.batch
    deg = original_degrees(g)
    cand = np.flatnonzero(deg >= 1)
    rng = np.random.default_rng(args.seed)
    sources = np.sort(rng.choice(cand, size=k, replace=False)).astype(np.uint32)
    bs = sw.BatchSession(g, args.algo, budget=budget, buffers=args.buffers, transfer=args.transfer, profile=True)
    t0 = time.perf_counter()
    unroll, tune_ms = bs.tune(sources)
    log(f"autotuner: unroll {unroll} (runs {', '.join(f'{t:.0f}' for t in tune_ms)} ms) in "
        f"{time.perf_counter() - t0:.1f} s")
    for i in range(args.warmup):
        _, st = bs.run(sources, rows_out=False)
        log(f"batch warmup {i}: {st.ms_total:.1f} ms")
    stats, wall = [], 0.0
    for i in range(args.steps):
        t0 = time.perf_counter()
        _, st = bs.run(sources, rows_out=False)
        wall += time.perf_counter() - t0
        stats.append(st)
        log(f"batch step {i}: {st.ms_total:.1f} ms e2e, {st.global_iterations} iters, H2D {st.h2d_bytes / 1e9:.3f} GB, "
            f"rounds {st.ms_rounds:.1f} ms, transfer {st.ms_transfer:.1f} ms")
    rows, _ = bs.run(sources, rows_out=True)     # one more (untimed) run hands the rows over
    bs.close()
    ms_batch = wall / args.steps * 1e3
    sess = sw.Session(g, args.algo, mode="async", budget=budget, buffers=args.buffers, transfer=args.transfer)
    sess.run(int(sources[0]), copy=False)
    mismatched = []
    t_seq, h2d_seq = 0.0, 0
    for i, s0 in enumerate(sources):
        t0 = time.perf_counter()
        v, st = sess.run(int(s0), copy=False)
        t_seq += time.perf_counter() - t0
        h2d_seq += st.h2d_bytes
        if not np.array_equal(v, rows[i]):
            mismatched.append(int(s0))
            log(f"source {s0}: batched row differs from the single-source run")
        if i % 8 == 7:
            log(f"sequential: {i + 1} of {k} runs, {t_seq:.2f} s")
    sess.close()
    ms_seq = t_seq * 1e3
    h2d = float(np.mean([s.h2d_bytes for s in stats]))
    xfer_ms = float(np.mean([s.ms_transfer for s in stats]))
    line = {
        "metric": METRIC,
        "value": round(k * g.m / (ms_batch / 1e3) / 1e9, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_batch, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic R-MAT built on the GPU in-process (no datasets)",
        "config": {
            "workload": f"{k}-source batched {args.algo}-async on weighted R-MAT scale {args.scale} ef "
                        f"{args.edge_factor}, sources uniform over out-degree >= 1 (seed {args.seed}), device edge "
                        f"budget {budget} B ({budget / GiB:g} GiB)",
            "vertices": g.n, "edges": g.m, "sources": k, "budget_bytes": budget, "buffers": args.buffers,
            "transfer": args.transfer, "unroll": unroll, "vertex_order": args.order,
        },
        "batched": {"ms_per_run": round(ms_batch, 3), "global_iterations": stats[-1].global_iterations,
                    "inner_rounds": stats[-1].inner_rounds, "rounds_launches": stats[-1].rounds_launches,
                    "work_edges": stats[-1].work_edges, "work_entries": stats[-1].work_entries,
                    "items_scanned": stats[-1].work_items, "updates": stats[-1].updates,
                    "lane_work": stats[-1].lane_work, "lane_entries": stats[-1].lane_entries,
                    "h2d_gb_per_run": round(h2d / 1e9, 4),
                    "h2d_gbps_achieved": round(h2d / (xfer_ms * 1e-3) / 1e9, 2) if xfer_ms > 0 else None,
                    "ms_breakdown": {"rounds": round(float(np.mean([s.ms_rounds for s in stats])), 3),
                                     "transfer": round(xfer_ms, 3),
                                     "generate": round(float(np.mean([s.ms_generate for s in stats])), 3)}},
        "sequential": {"ms_total": round(ms_seq, 3), "ms_per_source": round(ms_seq / k, 3),
                       "h2d_gb_total": round(h2d_seq / 1e9, 3), "gteps": round(k * g.m / t_seq / 1e9, 4)},
        "batched_speedup_vs_sequential": round(ms_seq / ms_batch, 3),
        "same_result_batched_vs_sequential": not mismatched,
        "h2d_gbps_peak": round(peak, 2),
        "roofline": batch_roofline(stats),
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return not mismatched


def batch_roofline(stats):
    """Batched rounds kernel, algorithmic bytes over all rounds: per relaxed
    edge s(E) once plus 8 B per source active at its vertex (destination
    distance 4 + atomic 4; lane_work is exactly the sum of out-degree x active
    sources over the pushing vertices); per pushing vertex its entry (id 4 +
    offset 8) and mask (8) plus 4 B per active source (its own distance;
    lane_entries)."""
    eb = stats[-1].edge_bytes
    ms = sum(s.ms_rounds for s in stats)
    launches = sum(s.rounds_launches for s in stats)
    work = sum(s.work_edges for s in stats)
    lanes = sum(s.lane_work for s in stats)
    lane_ents = sum(s.lane_entries for s in stats)
    ents = sum(s.work_entries for s in stats)
    alg = eb * work + 8 * lanes + 20 * ents + 4 * lane_ents
    achieved = alg / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"bound": "hbm", "kernel": "rounds_batch_kernel", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "launches": launches,
            "avg_launch_us": round(ms / max(launches, 1) * 1e3, 2),
            "algorithmic_bytes_per_launch": int(alg / max(launches, 1)), "edge_bytes": eb}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c5", choices=sorted(CONFIGS))
    ap.add_argument("--scale", type=int)
    ap.add_argument("--edge-factor", type=int)
    ap.add_argument("--algo")
    ap.add_argument("--mode", default="async")
    ap.add_argument("--budget")
    ap.add_argument("--buffers", type=int, default=2)
    ap.add_argument("--transfer", default="auto")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--variants", default="plain,rule,network,learned",
                    help="comma-separated subset of plain, rule, network, learned")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--batch", type=int)
    ap.add_argument("--order", default="in-degree", choices=["in-degree", "natural"],
                    help="vertex numbering the runs use: renumbered by descending in-degree on the GPU after the "
                         "build (results still in the original numbering), or the generator's own")
    ap.add_argument("--blocks", type=int, default=1,
                    help="1: build the block-compressed rows after the build (and renumbering) and move whole id "
                         "ranges in them; 0: plain edges only")
    ap.add_argument("--probe-copy-mb", type=int, default=0,
                    help="launch the plain copy kernel once on this many MiB (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): the multi-GPU path (disjoint "
              "partitions per GPU with a per-iteration all-reduce of vertex values, or batched sources split "
              "across GPUs) is not built", file=sys.stderr)
        sys.exit(2)
    preset = CONFIGS[args.config]
    for k, v in preset.items():
        if getattr(args, k, None) is None:
            setattr(args, k, v)
    args.weighted = preset["weighted"] if args.algo == preset["algo"] else args.algo in ("sssp", "sswp")
    args.symmetrize = args.algo == "cc"
    if args.algo in ("sssp", "sswp"):
        args.weighted = True

    import numpy as np

    import subway_amd as sw
    from subway_amd import policy

    if sw.device_count() < 1:
        log("no GPU visible")
        sys.exit(1)
    budget = parse_size(args.budget)
    t0 = time.perf_counter()
    g = sw.Graph.rmat(args.scale, args.edge_factor, seed=args.seed, weighted=args.weighted,
                      symmetrize=args.symmetrize)
    gen_s = time.perf_counter() - t0
    log(f"R-MAT scale {args.scale} ef {args.edge_factor}: {g.n} vertices, {g.m} edges built in {gen_s:.1f} s")
    src = g.max_degree_vertex
    reorder_s = None
    if args.order == "in-degree":
        t0 = time.perf_counter()
        g.reorder()
        reorder_s = time.perf_counter() - t0
        log(f"renumbered by descending in-degree in {reorder_s:.1f} s")
    blocks_s = block_bpe = None
    if args.blocks and not args.batch:
        t0 = time.perf_counter()
        # a graph at the size where pinned memory runs short holds only its
        # split arrays pinned; they move to ordinary memory before the block
        # words are pinned (compacted units are then gathered on the host)
        large = g.m * 5 >= 16 * GiB
        g.build_blocks(unpin_split=large and args.weighted)
        blocks_s = time.perf_counter() - t0
        _, words, _, _ = g.blocks_view()
        block_bpe = 4.0 * len(words) / max(g.m, 1)
        log(f"block layout built in {blocks_s:.1f} s: {block_bpe:.3f} bytes per edge")

    if args.probe_copy_mb:
        ms = sw.probe_copy(args.probe_copy_mb << 20)
        log(f"copy probe: {args.probe_copy_mb} MiB read + written in {ms:.3f} ms")
    peak = sw.h2d_peak(256 << 20, 10)
    log(f"pinned H2D peak {peak:.1f} GB/s")

    if args.batch:
        if not 1 <= args.batch <= 64 or args.algo not in ("sssp", "sswp", "bfs"):
            log("--batch takes 1..64 sources of sssp, sswp or bfs")
            sys.exit(2)
        sys.exit(0 if batch_bench(sw, np, g, args, budget, peak) else 1)

    kw = dict(mode=args.mode, budget=budget, buffers=args.buffers, transfer=args.transfer, profile=True)
    if args.algo == "pr":
        kw["threshold"] = 1e-6
    params = policy.default_params()
    names = [n for n in args.variants.split(",") if n]
    if not names or any(n not in VARIANTS for n in names):
        log(f"--variants takes a subset of {', '.join(VARIANTS)}")
        sys.exit(2)
    results = measure(sw, np, g, args, src, kw, names, params)
    head = "learned" if "learned" in results else names[-1]
    out, stats, vals = results[head]
    checks = {}
    for name in names:
        if name == head:
            continue
        a, b = results[name][2], vals
        checks[f"same_result_{name}_vs_{head}"] = (bool(np.array_equal(a, b)) if args.algo != "pr" else bool(
            np.max(np.abs(a.astype(np.float64) - b) / np.maximum(np.abs(a.astype(np.float64)), 1e-30)) < 1e-4))
    for name in names:
        log(f"{name}: {results[name][0]['ms_per_step']:.2f} ms per run")

    edge_bytes = stats[-1].edge_bytes
    # every correct schedule of an async run moves the edges of every vertex
    # the first iteration finds labelled (all of them, K8) and that then has
    # a finite value: at least the edges out of reachable vertices, once
    bound = None
    if args.algo in ("sssp", "bfs", "sswp"):
        reach = (vals != 4294967294) if args.algo != "sswp" else (vals != 0)
        deg = original_degrees(g)
        moved = int(deg[reach].sum(dtype=np.uint64))
        moved_bytes = moved * edge_bytes
        bv = g.blocks_view() if args.blocks else None
        if bv is not None:
            # the rows of the reached vertices in the block layout
            rowb = np.diff(bv[0]).astype(np.int64) * 4
            o = g.order()
            if o is not None:
                rowb = rowb[o[0]]
            moved_bytes = min(moved_bytes, int(rowb[reach].sum()))
        bound_ms = moved_bytes / (peak * 1e9) * 1e3
        bound = {"edges_out_of_reached_vertices": moved, "fraction_of_edges": round(moved / g.m, 4),
                 "bytes_per_edge": round(moved_bytes / max(moved, 1), 4), "gb": round(moved_bytes / 1e9, 3),
                 "ms_at_pinned_peak": round(bound_ms, 2),
                 "best_possible_speedup_over_plain": (round(results["plain"][0]["ms_per_step"] / bound_ms, 3)
                                                      if "plain" in results else None)}

    cpu = None
    if args.cpu_baseline:
        cpu = cpu_baseline(np, g, args, src, vals)
        checks["same_result_cpu_baseline_vs_gpu"] = cpu["same_result_as_gpu"]

    speed = {f"{n}_speedup_vs_plain": round(results["plain"][0]["ms_per_step"] / results[n][0]["ms_per_step"], 4)
             for n in names if n != "plain" and "plain" in results}
    line = {
        "metric": METRIC,
        "value": out["gteps"],
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": out["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32" if args.algo != "pr" else "f32",
        "data": "synthetic R-MAT built on the GPU in-process into pinned host memory (no datasets)",
        "config": {
            "workload": workload_name(args, budget),
            "name": args.config,
            "vertices": g.n, "edges": g.m, "source": src if args.algo not in ("cc", "pr") else None,
            "budget_bytes": budget, "buffers": args.buffers, "transfer": args.transfer, "policy": head,
            "scheduler_network": "subway_amd/data/scheduler_mlp.f32",
            "graph_build_s": round(gen_s, 2),
            "vertex_order": args.order,
            "reorder_s": round(reorder_s, 2) if reorder_s is not None else None,
            "blocks": bool(blocks_s is not None),
            "blocks_s": round(blocks_s, 2) if blocks_s is not None else None,
            "block_bytes_per_edge": round(block_bpe, 4) if block_bpe is not None else None,
        },
        "variants": {k: v[0] for k, v in results.items()},
        **speed,
        "learned_speedup_vs_plain": speed.get("learned_speedup_vs_plain"),
        "checks": checks,
        "network_prediction": prediction_error(stats[-1].trace) if head in ("learned", "network") else None,
        "transfer_lower_bound": bound,
        "h2d_gb_per_step": out["h2d_gb_per_step"],
        "h2d_gbps_achieved": out["h2d_gbps_achieved"],
        "h2d_gbps_peak": round(peak, 2),
        "h2d_frac": round(out["h2d_gbps_achieved"] / peak, 3) if out["h2d_gbps_achieved"] and peak else None,
        "roofline": roofline(stats, args),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    failed = [k for k, v in checks.items() if not v]
    if failed:
        log(f"FAILED checks: {', '.join(failed)}")
        sys.exit(1)


if __name__ == "__main__":
    main()
