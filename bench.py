"""Benchmark of the out-of-GPU-memory pipeline on one MI355X.

Workload (default): SSSP-async on a weighted R-MAT graph, scale 24
(16,777,216 vertices, ~268M edges, a=0.57 b=0.19 c=0.19, weights U[1,64],
self-loops dropped), from the highest out-degree vertex, with the device
edge budget capped at 1 GiB so every global iteration goes through
subgraph generation, partitioning and host-to-device transfer.

A step is one full run from the initial state to the fixpoint (first
generate .. values in the session's host result array), with the graph
resident in pinned host memory.  value = input edges / step time (GTEPS),
wall clock over the timed steps.  The same workload runs under the plain
Subway policy and under the learned scheduler (the network shipped in
subway_amd/data), each in its own session; the values of the two must be
identical.  Each timed step records HIP events around every rounds-kernel
launch (on its stream) and every transfer, from which the roofline entry and
the achieved H2D rate are computed.  The CPU baseline is the OpenMP build of
the same update rules on the same CSR, timed once on this host.

With --batch K (K <= 64) the workload is K-source batched SSSP instead: K
sources drawn uniformly (seeded) from the vertices with out-degree >= 1
advance together, one cooperative kernel per transferred partition, after
the autotuner has picked the kernel's unroll; a step is one batched run and
value = K * input edges / step time.  The same K sources then run one after
another as single-source SSSP-async sessions (plain policy, same budget);
the rows must be identical and the line reports the speedup.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "end-to-end GTEPS + H2D GB moved; learned vs plain policy, same exact result"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse_size(s):
    mul = {"K": 1 << 10, "M": 1 << 20, "G": 1 << 30}
    s = str(s)
    return int(float(s[:-1]) * mul[s[-1].upper()]) if s[-1].upper() in mul else int(s)


def profiled_traffic(kernel, workload_key):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE passes, each scaled
    by the copy-kernel calibration it records), or None."""
    path = os.path.join(ROOT, "profiles", "sssp_async_rmat24_hbm_counters.txt")
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if parts[:2] == ["traffic_bytes_per_launch", kernel] and " ".join(parts[2:-1]) == workload_key:
                    return int(parts[-1])
    except OSError:
        return None
    return None


def measure(sw, np, g, args, src, kw, names, params):
    """Warm-up and timed steps of each policy, interleaved step by step so
    that both see the same machine state; each run is timed on its own.
    Returns {name: (summary, stats, values of the last step)}."""
    sessions = {}
    for name in names:
        t0 = time.perf_counter()
        sessions[name] = sw.Session(g, args.algo, policy=name, params=params if name == "learned" else None, **kw)
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
            vals[name] = np.array(v, copy=True)
            stats[name].append(st)
            log(f"{name} step {i}: {st.ms_total:.1f} ms e2e, {st.global_iterations} iters, {st.partitions} partitions, "
                f"H2D {st.h2d_bytes / 1e9:.3f} GB, rounds {st.ms_rounds:.1f} ms ({st.rounds_launches} launches), "
                f"transfer {st.ms_transfer:.1f} ms, generate {st.ms_generate:.1f} ms, "
                f"mispredictions {st.mispredictions}, fallback at {st.fallback_iteration}")
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
            "h2d_gb_per_step": round(h2d / 1e9, 4),
            "h2d_mb_per_iteration": [round(r["h2d_bytes"] / 1e6, 2) for r in ss[-1].trace],
            "h2d_gbps_achieved": round(h2d / (xfer_ms * 1e-3) / 1e9, 2) if xfer_ms > 0 else None,
            "ms_breakdown": {"rounds": round(float(np.mean([s.ms_rounds for s in ss])), 3),
                             "transfer": round(xfer_ms, 3),
                             "generate": round(float(np.mean([s.ms_generate for s in ss])), 3)},
            "mispredictions": ss[-1].mispredictions,
            "fallback_iteration": ss[-1].fallback_iteration,
        }, ss, vals[name])
    for s in sessions.values():
        s.close()
    return out


def batch_bench(sw, np, g, args, budget, peak):
    """K-source batched SSSP against K sequential single-source runs."""
    k = args.batch
    deg = g.out_degree()
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
    for r in stats[-1].trace:
        log(f"  batch iter {r['iteration']}: active {r['active_nodes']} v / {r['active_edges']} e, "
            f"{r['partitions']} partitions, H2D {r['h2d_bytes'] / 1e6:.1f} MB, {r['ms']:.2f} ms")
    # one more (untimed) run hands the rows to Python for the comparison
    rows, _ = bs.run(sources, rows_out=True)
    bs.close()
    ms_batch = wall / args.steps * 1e3
    # k sequential single-source runs of the same sources
    sess = sw.Session(g, args.algo, mode="async", budget=budget, buffers=args.buffers, transfer=args.transfer)
    sess.run(int(sources[0]), copy=False)
    same = True
    t_seq = 0.0
    h2d_seq = 0
    for i, s0 in enumerate(sources):
        t0 = time.perf_counter()
        v, st = sess.run(int(s0), copy=False)
        t_seq += time.perf_counter() - t0
        h2d_seq += st.h2d_bytes
        if not np.array_equal(v, rows[i]):
            same = False
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
        "data": "synthetic R-MAT generated in-process (no datasets)",
        "config": {
            "workload": f"{k}-source batched {args.algo} on weighted R-MAT scale {args.scale} ef {args.edge_factor}, "
                        f"sources uniform over out-degree >= 1 (seed {args.seed}), device edge budget {budget} B",
            "vertices": g.n, "edges": g.m, "sources": k, "budget_bytes": budget, "buffers": args.buffers,
            "transfer": args.transfer, "unroll": unroll,
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
        "same_result_batched_vs_sequential": same,
        "h2d_gbps_peak": round(peak, 2),
        "roofline": batch_roofline(stats),
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def batch_roofline(stats):
    """Batched rounds kernel, algorithmic bytes over all rounds: per relaxed
    edge s(E) once plus 8 B per source active at its vertex (destination
    distance 4 + atomic 4; the lane_work counter is exactly the sum of
    out-degree x active sources over the pushing vertices); per pushing
    vertex its entry (id 4 + offset 8) and mask (8) plus 4 B per active
    source (its own distance; lane_entries)."""
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


def roofline(np, stats, args, budget):
    """Roofline of the dominant kernel.  Persistent async rounds: one
    rounds_kernel launch per partition visit does every inner round (frontier
    + advance).  Algorithmic bytes, summed over the rounds: label algorithms
    22 B per labelled active vertex (id 4 + offset 8 + degree 4 + label r/w 2
    + value 4), s(E) + 4 B per relaxed edge (edge as buffered + destination
    value), 9 B per successful update (atomic RMW 8 + label 1); PageRank
    (FP64 residual) 48 B per pushing vertex (id 4 + offset 8 + degree 4 +
    residual read and clear 16 + running-sum RMW 16) and s(E) + 16 B per
    edge (FP64 atomic add RMW).  The frontier pass over unlabelled items is
    not counted.  Partition visits run either as one persistent rounds_kernel
    launch or as stand-alone frontier + advance launch pairs (a round cap of
    one, launch-per-round mode, oversized vertices); the run's counters tell
    the pairs' share apart, and the kind with more kernel time is reported,
    priced with its own work only."""
    eb = stats[-1].edge_bytes
    rounds_ms = sum(s.ms_rounds for s in stats)
    rounds_launches = sum(s.rounds_launches for s in stats)
    pair_ms = sum(s.ms_advance + s.ms_frontier for s in stats)
    adv_launches = sum(s.advance_launches for s in stats)
    if rounds_launches and rounds_ms >= pair_ms:
        kern, k_ms, k_launches = "rounds_kernel", rounds_ms, rounds_launches
        work = sum(s.work_edges - s.pair_work_edges for s in stats)
        upd = sum(s.updates - s.pair_updates for s in stats)
        ents = sum(s.work_entries - s.pair_work_entries for s in stats)
    else:
        kern, k_ms, k_launches = "advance_kernel+frontier_kernel", pair_ms, adv_launches
        work = sum(s.pair_work_edges for s in stats)
        upd = sum(s.pair_updates for s in stats)
        ents = sum(s.pair_work_entries for s in stats)
    if args.algo == "pr":
        alg_bytes = 48 * ents + (eb + 16) * work
    else:
        alg_bytes = 22 * ents + (eb + 4) * work + 9 * upd
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    traffic = None
    if (kern == "rounds_kernel" and args.algo == "sssp" and args.mode == "async" and args.scale == 24
            and args.edge_factor == 16 and budget == 1 << 30 and args.buffers == 2):
        traffic = profiled_traffic(kern, "sssp-async rmat24 budget1GiB learned")
    return {
        "bound": "hbm",
        "kernel": kern,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "launches": k_launches,
        "avg_launch_us": round(k_ms / max(k_launches, 1) * 1e3, 2),
        "algorithmic_bytes_per_launch": int(alg_bytes / max(k_launches, 1)),
        "edge_bytes": eb,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--algo", default="sssp")
    ap.add_argument("--mode", default="async")
    ap.add_argument("--budget", default="1G")
    ap.add_argument("--buffers", type=int, default=2)
    ap.add_argument("--transfer", default="auto")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--policy", default="both", choices=["both", "plain", "learned"])
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--batch", type=int, default=0, help="K-source batched SSSP against K sequential runs")
    ap.add_argument("--probe-copy-mb", type=int, default=0,
                    help="launch the plain copy kernel once on this many MiB (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): the multi-GPU path (disjoint "
              "partitions per GPU with a per-iteration all-reduce of vertex values, or batched sources split "
              "across GPUs) is not built", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import subway_amd as sw
    from subway_amd import policy

    if sw.device_count() < 1:
        log("no GPU visible")
        sys.exit(1)
    weighted = args.algo in ("sssp", "sswp")
    budget = parse_size(args.budget)
    t0 = time.perf_counter()
    g = sw.Graph.rmat(args.scale, args.edge_factor, seed=args.seed, weighted=weighted,
                      symmetrize=(args.algo == "cc"))
    log(f"R-MAT scale {args.scale}: {g.n} vertices, {g.m} edges in {time.perf_counter() - t0:.1f} s")
    src = g.max_degree_vertex
    kw = dict(mode=args.mode, budget=budget, buffers=args.buffers, transfer=args.transfer, profile=True)
    if args.algo == "pr":
        kw["threshold"] = 1e-6

    if args.probe_copy_mb:
        ms = sw.probe_copy(args.probe_copy_mb << 20)
        log(f"copy probe: {args.probe_copy_mb} MiB read + written in {ms:.3f} ms")
    peak = sw.h2d_peak(256 << 20, 10)
    log(f"pinned H2D peak {peak:.1f} GB/s")

    if args.batch:
        if not 1 <= args.batch <= 64 or args.algo not in ("sssp", "sswp", "bfs"):
            log("--batch takes 1..64 sources of sssp, sswp or bfs")
            sys.exit(2)
        batch_bench(sw, np, g, args, budget, peak)
        return

    params = policy.default_params()
    names = ["plain", "learned"] if args.policy == "both" else [args.policy]
    results = measure(sw, np, g, args, src, kw, names, params)
    head = "learned" if "learned" in results else "plain"
    out, stats, vals = results[head]
    same = None
    if len(results) == 2:
        a, b = results["plain"][2], results["learned"][2]
        same = bool(np.array_equal(a, b)) if args.algo != "pr" else bool(
            np.max(np.abs(a - b) / np.maximum(np.abs(a), 1e-30)) < 1e-4)
        log(f"plain vs learned: {results['plain'][0]['ms_per_step']:.2f} vs {results['learned'][0]['ms_per_step']:.2f}"
            f" ms per run, same result: {same}")

    cpu = None
    if args.cpu_baseline:
        from tests import _oracle as O
        ptr, adj, w = g.csr()
        t = time.perf_counter()
        if args.algo == "pr":
            cval, _, _ = O.omp_pr(ptr, adj, 1e-6)
        else:
            cval, _, _ = O.omp_label(args.algo, ptr, adj, w, src)
        ct = time.perf_counter() - t
        if args.algo != "pr" and not np.array_equal(cval, vals):
            log("WARNING: CPU baseline and GPU results differ")
        cpu = {"value": g.m / ct / 1e9, "unit": "GTEPS", "cores": O.omp_threads(), "kind": "port",
               "sample": f"one full {args.algo} run (OpenMP frontier rounds, same CSR and source), {ct:.2f} s"}
        log(f"cpu baseline: {ct:.2f} s on {O.omp_threads()} threads")

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
        "data": "synthetic R-MAT generated in-process (no datasets)",
        "config": {
            "workload": f"{args.algo}-{args.mode} on weighted R-MAT scale {args.scale} ef {args.edge_factor}, "
                        f"source = max out-degree vertex, device edge budget {budget} B",
            "vertices": g.n, "edges": g.m, "source": src, "budget_bytes": budget, "buffers": args.buffers,
            "transfer": args.transfer, "policy": head,
            "scheduler_network": "subway_amd/data/scheduler_mlp.f32" if params is not None else None,
        },
        "policies": {k: v[0] for k, v in results.items()},
        "learned_trace": ([{"iteration": r["iteration"], "features": [round(f, 5) for f in r["features"]],
                            "heads": [round(h, 5) for h in r["heads"]], "theta": r["theta"],
                            "rounds_cap": r["rounds_cap"], "on_plain": r["on_plain"]}
                           for r in stats[-1].trace] if head == "learned" else None),
        "learned_speedup_vs_plain": (round(results["plain"][0]["ms_per_step"] / results["learned"][0]["ms_per_step"], 4)
                                     if len(results) == 2 else None),
        "same_result_learned_vs_plain": same,
        "h2d_gb_per_step": out["h2d_gb_per_step"],
        "h2d_gbps_achieved": out["h2d_gbps_achieved"],
        "h2d_gbps_peak": round(peak, 2),
        "h2d_frac": round(out["h2d_gbps_achieved"] / peak, 3) if out["h2d_gbps_achieved"] and peak else None,
        "roofline": roofline(np, stats, args, budget),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
