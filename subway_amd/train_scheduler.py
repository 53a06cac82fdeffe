"""Trains the learned transfer/async scheduler from traces the pipeline
records itself, on the GPU, with the library's own forward/backward/Adam
kernels.

    python -m subway_amd.train_scheduler --out subway_amd/data/scheduler_mlp.f32

The training set is the small-graph sweep (R-MAT scales 16-23, budgets of a
whole, a half and a sixth of the graph) plus the benchmark configurations
at their stated sizes (BFS-sync scale 16 with no cap; SSSP-/SSWP-async scale
24 under 1 GiB; CC-async on the symmetrised scale 25 and PR-async scale 25,
2 GiB; SSSP-/BFS-async scale 26 under 4 GiB).  A sweep can be split over
several processes (--set, --only i:j, --dump) and trained from the dumps
(--train-from).  The graphs are renumbered by descending in-degree and
carry their block-compressed rows, as the benchmark's do (--order natural
keeps the generator's ids, --blocks 0 the plain transfers).

For every training configuration (algorithm, R-MAT graph, device budget) the
async run is repeated under a grid of fixed schedules -- whole-range density
threshold theta x inner-round cap -- with the learned policy's unit builder
and no network, so every iteration's 12 features are recorded by the policy
kernel; the plain policy runs beside them for reference.  All sessions of a
configuration are timed round-robin in random order.  Each decision of a configuration is labelled by its own best value (median
end-to-end time, the other decision at its best; near-ties within 2% go to
a fixed preference order) for every iteration of that schedule's run; head 0
(log2 of the next iteration's active-edge fraction) is labelled from every
run.  Results are bit-exact under every schedule, so the
sweep only measures time.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

import subway_amd as sw
from subway_amd import policy

THETAS = (2.0, 0.8, 0.5, 0.25)     # 2.0: no range moves whole (the plain transfer choice)
CAPS = (4096, 64, 8, 1)
PLAIN_LIKE = (2.0, 4096)
# near-ties (within 2%) go to the earlier entry
THETA_PREF = (0.8, 0.5, 2.0, 0.25)
CAP_PREF = (64, 8, 1, 4096)


# the benchmark configurations: (algo, scale, edge factor, seed, budget bytes, symmetrise)
NORTH = [("bfs", 16, 16, 42, 0, False), ("sssp", 24, 16, 42, 1 << 30, False), ("sswp", 24, 16, 42, 1 << 30, False),
         ("cc", 25, 16, 42, 2 << 30, True), ("pr", 25, 16, 42, 2 << 30, False),
         ("sssp", 26, 16, 42, 4 << 30, False), ("bfs", 26, 16, 42, 4 << 30, False),
         # graphs of the benchmark's size on another seed: the next-volume head
         # sees iterations of 8.6 G-edge graphs
         ("sssp", 28, 32, 7, 16 << 30, False), ("bfs", 28, 32, 7, 16 << 30, False),
         ("sswp", 27, 32, 7, 8 << 30, False)]


def configs(quick: bool, which: str = "small"):
    """Training configurations as (algo, scale, edge factor, seed, budget
    divisor, budget bytes, symmetrise): the small sweep's device budget holds
    1/div of the graph's edges (two buffers); the benchmark configurations
    give their budget in bytes (0: no cap)."""
    out = []
    if which in ("north", "all"):
        out += [(a, s, ef, seed, 0, b, sym) for a, s, ef, seed, b, sym in NORTH]
    if which == "north":
        return out
    for algo in ("sssp", "sswp", "bfs", "pr"):
        if quick:
            scales = (15, 16)
        elif algo == "pr":
            scales = (16, 18, 20)
        elif algo in ("sssp", "bfs"):
            scales = (16, 17, 18, 19, 20, 21, 22, 23)
        else:
            scales = (16, 17, 18, 19, 20, 21)
        for scale in scales:
            for seed in ((1,) if quick or scale >= 20 else (1, 2)):
                for div in ((2, 6) if quick else (1, 2, 6)):
                    out.append((algo, scale, 16, seed, div, 0, False))
    return out


def make_graph(algo, scale, seed, edge_factor=16, symmetrize=False, order="in-degree", blocks=True):
    """The R-MAT graph of a configuration, renumbered by descending in-degree
    unless order is "natural" (the benchmark renumbers its graphs, and the
    schedule's id ranges see the numbering), with its block-compressed rows
    unless blocks is False (the benchmark moves whole ranges in them, which
    changes what moving a range whole costs)."""
    g = sw.Graph.rmat(scale, edge_factor, seed=seed, weighted=algo in ("sssp", "sswp"), symmetrize=symmetrize)
    if order == "in-degree":
        g.reorder()
    if blocks:
        g.build_blocks()
    return g


def run_config(g, algo, div, reps, log, seed=0, budget_bytes=None):
    """Median time and the trace of each fixed schedule, and of the plain
    policy.  All sessions exist at once and are timed round-robin in a fresh
    random order each round, so warm-up and clock drift touch every schedule
    alike."""
    edge_bytes = 4    # weighted R-MAT at these scales packs into one word; unweighted edges are 4 B
    budget = budget_bytes if budget_bytes is not None else max(64 << 10, (g.m * edge_bytes) // div)
    src = g.max_degree_vertex if algo != "pr" else 0
    kw = dict(mode="async", budget=budget)
    if algo == "pr":
        kw["threshold"] = 1e-6
    keys = [(th, cap) for th in THETAS for cap in CAPS] + ["plain"]

    def open_session(k):
        if k == "plain":
            return sw.Session(g, algo, policy="plain", **kw)
        return sw.Session(g, algo, policy="learned", theta=k[0], max_rounds=k[1], **kw)

    ref = None
    times = {k: [] for k in keys}
    traces = {}
    rng = np.random.default_rng(seed)

    def one_run(s, k, r):
        nonlocal ref
        vals, st = s.run(src)
        if st.ms_total > 1000:
            log(f"    round {r}, schedule {k}: {st.ms_total:.0f} ms")   # long runs report as they go
        if r < 2:
            # warm-up rounds; the values must not depend on the schedule
            if ref is None:
                ref = np.array(vals, copy=True)
            elif algo != "pr" and not np.array_equal(ref, vals):
                raise RuntimeError(f"{algo}: schedule {k} changed the result")
            return
        times[k].append(st.ms_total)
        traces[k] = st.trace

    if budget * len(keys) > (96 << 30):
        # the device does not hold every session's buffers at once: one
        # schedule at a time, in a random order
        order = list(keys)
        rng.shuffle(order)
        for k in order:
            s = open_session(k)
            for r in range(reps + 2):
                one_run(s, k, r)
            s.close()
    else:
        sess = {k: open_session(k) for k in keys}
        for r in range(reps + 2):
            order = list(keys)
            rng.shuffle(order)
            for k in order:
                one_run(sess[k], k, r)
        for s in sess.values():
            s.close()
    res = {k: (float(np.median(times[k])), [{"features": list(r["features"])} for r in traces[k]])
           for k in keys if k != "plain"}
    plain_ms = float(np.median(times["plain"]))
    # each decision labelled by its own best (the other decision at its best)
    t_theta = {th: min(res[(th, c)][0] for c in CAPS) for th in THETAS}
    t_cap = {c: min(res[(th, c)][0] for th in THETAS) for c in CAPS}
    th_best = next(th for th in THETA_PREF if t_theta[th] <= 1.02 * min(t_theta.values()))
    cap_best = next(c for c in CAP_PREF if t_cap[c] <= 1.02 * min(t_cap.values()))
    best = (th_best, cap_best)
    log(f"  {algo} n={g.n} m={g.m} budget={budget}: plain {plain_ms:.2f} ms, plain-like {res[PLAIN_LIKE][0]:.2f} ms, "
        f"label {best} {res[best][0]:.2f} ms, fastest {min(r[0] for r in res.values()):.2f} ms")
    return res, best, plain_ms


def dataset(results):
    xs, ys, ms = [], [], []
    for res, best in results:
        for sched, (_, trace) in res.items():
            for t, rec in enumerate(trace):
                nxt = trace[t + 1]["features"][2] if t + 1 < len(trace) else 0.0
                y = [max(np.log2(nxt + 2.0 ** -30), -30.0) / 32.0, best[0], np.log2(best[1]) / 16.0]
                m = [1.0, 1.0, 1.0] if sched == best else [1.0, 0.0, 0.0]
                xs.append(rec["features"])
                ys.append(y)
                ms.append(m)
    return np.array(xs, np.float32), np.array(ys, np.float32), np.array(ms, np.float32)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "data", "scheduler_mlp.f32"))
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--save-data", default="", help="also write the training samples to this .npz")
    ap.add_argument("--set", default="all", choices=["small", "north", "all"])
    ap.add_argument("--only", default="", help="i:j, a slice of the configuration list")
    ap.add_argument("--dump", default="", help="write the sweep's results (times, traces) to this JSON and stop")
    ap.add_argument("--train-from", nargs="*", default=[], help="train from sweep dumps instead of sweeping")
    ap.add_argument("--order", default="in-degree", choices=["in-degree", "natural"],
                    help="vertex numbering of the sweep's graphs")
    ap.add_argument("--blocks", type=int, default=1, help="1: the sweep's graphs move whole ranges as blocks")
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[train] {msg}", file=sys.stderr, flush=True)

    t0 = time.time()
    results, summary = [], []
    if args.train_from:
        for path in args.train_from:
            with open(path) as f:
                for c in json.load(f):
                    res = {tuple(k): (v[0], v[1]) for k, v in c["results"]}
                    results.append((res, tuple(c["summary"]["best"])))
                    summary.append(c["summary"])
    else:
        todo = configs(args.quick, args.set)
        if args.only:
            i, j = (int(x) if x else None for x in args.only.split(":"))
            todo = todo[i:j]
        dumped = []
        for algo, scale, ef, seed, div, budget, sym in todo:
            g = make_graph(algo, scale, seed, ef, sym, args.order, bool(args.blocks))
            big = g.m >= (1 << 28)
            res, best, plain_ms = run_config(g, algo, div, 1 if big else args.reps, log, seed=len(results),
                                             budget_bytes=budget if not div else None)
            g.free()
            results.append((res, best))
            s = {"algo": algo, "scale": scale, "edge_factor": ef, "seed": seed, "budget_divisor": div,
                 "budget_bytes": budget, "symmetrize": sym, "order": args.order, "blocks": bool(args.blocks),
                 "plain_ms": plain_ms,
                 "best": list(best),
                 "ms": {f"{k[0]}/{k[1]}": v[0] for k, v in res.items()}}
            summary.append(s)
            dumped.append({"summary": s, "results": [[list(k), [v[0], v[1]]] for k, v in res.items()]})
        if args.dump:
            with open(args.dump, "w") as f:
                json.dump(dumped, f)
            log(f"wrote {len(dumped)} configurations to {args.dump} in {time.time() - t0:.0f} s")
            return
    x, y, mask = dataset(results)
    log(f"{len(x)} samples from {len(results)} configurations in {time.time() - t0:.0f} s")
    if args.save_data:
        np.savez(args.save_data, x=x, y=y, mask=mask)
    mean, std = x.mean(0), x.std(0)
    p0 = policy.init_params(args.seed, mean, std)
    params, losses = policy.train(p0, x, y, mask, steps=args.steps, lr=args.lr)
    log(f"loss {losses[0]:.5f} -> {losses[-1]:.5f}")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    policy.save(args.out, params)
    meta = {"samples": int(len(x)), "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
            "steps": args.steps, "lr": args.lr, "thetas": THETAS, "caps": CAPS, "configs": summary}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
