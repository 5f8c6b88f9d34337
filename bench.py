"""Exact group Steiner tree queries/s on one MI355X: learned upper bound vs plain DPBF.

Workload (one step = one exact GST query, solved through the C-ABI with the graph resident in HBM):
  Barabasi-Albert V=200,000, m=5 (~1.0M undirected edges), int32 weights uniform in [1,1000],
  g=6 groups of sizes uniform in [10,200]; query q uses group seed 1000+q (64 distinct queries,
  cycled to --steps).

Headline `value`: queries/s of the batched learned-bound solve (gst_solve_batch_learned): the
trained network (gpu4gst_code_amd/models/mpnn_ba50k.npy) scores vertices, the DPBF optimum of
the top-3% induced subgraph bounds the full DPBF, subgraph and bounded full solves of up to 64
queries share the batched engine's launches.  Every cost is checked against the plain DPBF of the
same query, run one query at a time right after (reported as `plain_dpbf`).

Printed: one JSON line with the metric, the speedup over plain DPBF, the roofline of the dominant
kernel of the timed run (batched grow: algorithmic bytes / HIP-event kernel time vs 8 TB/s) and of
the batched merge, the plain path's grow/merge rooflines, state relaxations/s, and a CPU baseline
(the oracle's label-setting DPBF on the host cores over a bounded sample of the same queries).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="ba200k", choices=["ba200k", "c4", "rmat22", "rmat21"],
                    help="ba200k: the headline (default); c4: 256 mixed queries on the BA graph; rmat22: float "
                         "R-MAT scale 22, g=7; rmat21: learned bound on int R-MAT scale 21, g=8")
    ap.add_argument("--steps", type=int, default=256, help="queries in the timed run (the 64 queries cycled)")
    ap.add_argument("--warmup", type=int, default=128,
                    help="queries solved before the timed run (enough to fill the engine's slots once)")
    ap.add_argument("--n", type=int, default=200_000)
    ap.add_argument("--groups", type=int, default=6)
    ap.add_argument("--frac", type=float, default=0.05, help="top-scored fraction of vertices kept in the subgraph")
    ap.add_argument("--slots", type=int, default=128, help="queries in flight in the batched engine")
    ap.add_argument("--model", default=os.path.join(ROOT, "gpu4gst_code_amd", "models", "mpnn_ba50k.npy"))
    ap.add_argument("--autotune", type=int, default=1, help="1: pick launch knobs with gst_autotune before timing")
    ap.add_argument("--rounds-per-step", type=int, default=0, help="engine rounds per step (0: tuned / default)")
    ap.add_argument("--cpu-sample", type=int, default=16, help="queries in the CPU-baseline sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="MiB copied by one plain device copy before the run (counter calibration)")
    return ap.parse_args()


def gbs(nbytes, ms):
    return nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent queries would shard one-per-GPU with no communication)", file=sys.stderr)
        sys.exit(2)
    import gpu4gst_code_amd as gst

    if args.workload != "ba200k":
        return other_workload(args, gst)
    G = gst.gen_barabasi_albert(args.n, 5, 1, 1000, seed=1)
    queries = [gst.gen_groups(G.n, args.groups, 10, 200, seed=1000 + q) for q in range(64)]
    timed_q = [queries[i % 64] for i in range(args.steps)]
    solver = gst.Solver(0)
    dg = solver.upload(G)
    if not os.path.exists(args.model):
        raise SystemExit(f"bench.py: trained model {args.model} is missing")
    model = gst.Model(solver)
    model.load(args.model)

    if args.calibrate_copy > 0:
        import ctypes
        calib = ctypes.CDLL(os.path.join(ROOT, "gpu4gst_code_amd", "libgst_calib.so"))
        calib.gst_calib_copy.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int]
        assert calib.gst_calib_copy(0, args.calibrate_copy << 20, 1) == 0

    # launch knobs for this graph (bounded one-query solves and the batched engine), picked by the
    # project's autotuner on 16 of the queries, the engine timed as in the timed run (kernel by kernel
    # with events); plain solves keep the built-in defaults
    tune = None
    if args.autotune:
        t_tune = time.perf_counter()
        tune = solver.autotune(dg, queries[:16], bounded=True, batch=True, eager=True)
        tune["seconds"] = round(time.perf_counter() - t_tune, 2)

    # warm-up: engine slots, level tables, captured steps, the model's buffers, the plain path
    if args.warmup > 0:
        solver.solve_batch_learned(dg, [queries[i % 64] for i in range(args.warmup)], model, args.frac,
                                   slots=args.slots, rounds_per_step=args.rounds_per_step, time_kernels=True)
        for i in range(min(args.warmup, 4)):
            solver.solve(dg, queries[i % 64])

    # ---- timed: batched learned-bound solve (each call ends with a stream synchronisation) ----
    t0 = time.perf_counter()
    res, bst = solver.solve_batch_learned(dg, timed_q, model, args.frac, slots=args.slots,
                                          rounds_per_step=args.rounds_per_step, time_kernels=True)
    wall = time.perf_counter() - t0
    qps = args.steps / wall

    # ---- plain DPBF, one query at a time, on the 64 distinct queries ----
    t1 = time.perf_counter()
    plain = [solver.solve(dg, q) for q in queries]
    plain_wall = time.perf_counter() - t1
    plain_qps = len(queries) / plain_wall

    # the same queries with the automatic spider bound alone (no network), batched the same way
    t3 = time.perf_counter()
    sres, _ = solver.solve_batch(dg, timed_q, slots=args.slots, rounds_per_step=args.rounds_per_step, auto_bound=True,
                                 time_kernels=True)
    spider_qps = args.steps / (time.perf_counter() - t3)
    for i, r in enumerate(sres):
        assert r.cost == plain[i % 64].cost, (i, r.cost, plain[i % 64].cost)

    # exactness: every learned-bound cost equals plain DPBF; every tree is valid at that cost
    for i, r in enumerate(res):
        want = plain[i % 64].cost
        assert r.cost == want, (i, r.cost, want)
        if i < 64:
            tc = gst.validate_tree(G, timed_q[i], r.eu, r.ev, r.root)
            assert tc == want, (i, tc, want)
    fallback = sum(r.fell_back for r in res) / len(res)
    ratios = [r.stats["ub"] / plain[i % 64].cost for i, r in enumerate(res) if r.stats["ub"] > 0 and plain[i % 64].cost > 0]

    # rooflines: batched engine (timed run) and plain path
    b_grow_gbs = gbs(bst["grow_bytes"], bst["t_grow_ms"])
    b_merge_gbs = gbs(bst["merge_bytes"], bst["t_merge_ms"])
    p_grow_ms = sum(r.stats["t_grow_ms"] for r in plain)
    p_merge_ms = sum(r.stats["t_merge_ms"] for r in plain)
    p_grow_bytes = sum(r.stats["grow_bytes"] for r in plain)
    p_merge_bytes = sum(r.stats["merge_bytes"] for r in plain)
    p_relax = sum(r.stats["grow_relax"] for r in plain)
    p_launches = sum(r.stats["grow_launches"] for r in plain)

    cpu = None
    if args.cpu_sample > 0:
        import oracle
        sample = queries[: args.cpu_sample]
        nth = min(args.cpu_threads, len(sample), os.cpu_count() or 1)
        t2 = time.perf_counter()
        costs = oracle.dpbf_batch(G, sample, nthreads=nth)
        cw = time.perf_counter() - t2
        for i, c in enumerate(costs):
            assert c == plain[i].cost, (i, c, plain[i].cost)
        cpu = {"value": round(len(sample) / cw, 4), "unit": "queries/s", "cores": nth, "kind": "port",
               "sample": f"{len(sample)} of the 64 queries, label-setting DPBF (oracle/dpbf_oracle.c), "
                         f"{nth} threads, {cw:.1f} s wall"}
        # PrunedDP++-style bounded label-setting search (one-label and tour lower bounds, grow <= U/2)
        # with the spider bound computed on the host, one thread, on 4 of the queries
        t4 = time.perf_counter()
        for qi in range(4):
            d = oracle.group_distances(G, queries[qi])
            c, _ = oracle.dpbf_pruned(G, queries[qi], float(np.min(d.sum(axis=0))), 0.5, 2)
            assert c == plain[qi].cost, (qi, c, plain[qi].cost)
        pw = time.perf_counter() - t4
        cpu["pruned"] = {"value": round(4 / pw, 4), "unit": "queries/s", "cores": 1, "kind": "port",
                         "sample": f"4 of the 64 queries, bounded label-setting DPBF (oracle orc_dpbf_pruned, "
                                   f"spider bound, one-label and tour LBs, half-bound grow), {pw:.1f} s wall"}

    glaunch = max(int(bst["grow_launches"]), 1)
    # HBM traffic per grow launch from the committed rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE,
    # calibrated): their bytes per dispatch over their algorithmic bytes per dispatch, applied to this
    # run's algorithmic bytes per launch (under the profiler the host is slower, so its steps hold
    # fewer queries and its launches are smaller than here)
    traffic, t_ratio = None, None
    tpath = os.path.join(ROOT, "profiles", "dpbf_learned_batch_ba200k_g6_traffic.json")
    if os.path.exists(tpath):
        tg = json.load(open(tpath)).get("k_batch_grow", {})
        if tg.get("traffic_bytes_per_dispatch") and tg.get("algorithmic_bytes_per_dispatch"):
            t_ratio = tg["traffic_bytes_per_dispatch"] / tg["algorithmic_bytes_per_dispatch"]
            traffic = int(t_ratio * bst["grow_bytes"] / glaunch)
    mlaunch = max(int(bst["merge_launches"]), 1)
    line = {
        "metric": "exact-cost GST queries/s (learned-UB batched DPBF vs plain DPBF, BA V=200k g=6)",
        "value": round(qps, 3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded BA graph and groups)",
        "config": {"workload": f"BA V={G.n} m=5 E={G.m // 2} int32 w[1,1000], g={args.groups} sizes 10-200, "
                               f"{args.steps} queries (64 distinct), learned bound top {args.frac:.0%} subgraph, "
                               f"batched engine {args.slots} slots",
                   "mode": "learned-UB batched DPBF"},
        "plain_dpbf": {"value": round(plain_qps, 3), "unit": "queries/s", "mode": "plain DPBF, one query at a time"},
        "speedup_vs_plain": round(qps / plain_qps, 3),
        "spider_bound_batched": {"value": round(spider_qps, 3), "unit": "queries/s",
                                 "mode": "automatic spider bound only (no network), same batched engine"},
        "learned": {"frac": args.frac, "fallback_rate": round(fallback, 4),
                    "ub_over_opt_mean": round(float(np.mean(ratios)), 5) if ratios else None,
                    "relax_vs_plain": round(sum(r.stats["grow_relax"] for r in res[:64]) / max(p_relax, 1), 5)},
        "latency_ms": {"p50": round(bst["latency_p50_ms"], 3), "p99": round(bst["latency_p99_ms"], 3)},
        "autotune": tune,
        "state_relax_per_s": round(p_relax / (p_grow_ms * 1e-3), 1) if p_grow_ms > 0 else None,
        "roofline": {"bound": "hbm", "achieved": round(b_grow_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(b_grow_gbs / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_over_algorithmic": round(t_ratio, 2) if t_ratio else None,
                     "kernel": "k_batch_grow", "launches": glaunch,
                     "bytes_per_launch": round(bst["grow_bytes"] / glaunch, 1),
                     "ms_per_launch": round(bst["t_grow_ms"] / glaunch, 5)},
        "merge_roofline": {"bound": "hbm", "achieved": round(b_merge_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(b_merge_gbs / HBM_PEAK_GBS, 5), "kernel": "k_batch_merge",
                           "launches": mlaunch, "ms_per_launch": round(bst["t_merge_ms"] / mlaunch, 5)},
        "plain_roofline": {"grow": {"kernel": "k_grow_round", "achieved": round(gbs(p_grow_bytes, p_grow_ms), 2),
                                    "frac": round(gbs(p_grow_bytes, p_grow_ms) / HBM_PEAK_GBS, 5),
                                    "launches": p_launches,
                                    "ms_per_launch": round(p_grow_ms / max(p_launches, 1), 5)},
                           "merge": {"kernel": "k_merge", "achieved": round(gbs(p_merge_bytes, p_merge_ms), 2),
                                     "frac": round(gbs(p_merge_bytes, p_merge_ms) / HBM_PEAK_GBS, 5)}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    model.free()
    solver.close()


def _rates(results, key_ms):
    return sum(r.stats[key_ms] for r in results)


def other_workload(args, gst):
    """The other measured configurations (one JSON line each; not the headline)."""
    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)  # noqa: E731
    solver = gst.Solver(0)
    model = gst.Model(solver)
    model.load(args.model)
    out = {"n_gpus": 1, "data": "synthetic (seeded generators)", "higher_is_better": True, "vs_baseline": None}
    if args.workload == "c4":
        # 256 queries on the BA V=200k graph, g uniform in 3..8, group sizes uniform in 5..500
        G = gst.gen_barabasi_albert(200_000, 5, 1, 1000, seed=1)
        dg = solver.upload(G)
        rng = np.random.default_rng(5)
        qs = [gst.gen_groups(G.n, int(rng.integers(3, 9)), 5, 500, seed=5000 + q) for q in range(256)]
        log("graph and queries ready")
        solver.autotune(dg, [q for q in qs if q.g == 6][:8], bounded=True, batch=True)
        solver.solve_batch(dg, qs[:16], slots=args.slots)
        t = time.perf_counter(); pb, pst = solver.solve_batch(dg, qs, slots=args.slots); t_pb = time.perf_counter() - t
        log(f"plain batched {len(qs) / t_pb:.1f} q/s")
        t = time.perf_counter(); sb, sst = solver.solve_batch(dg, qs, slots=args.slots, auto_bound=True)
        t_sb = time.perf_counter() - t
        solver.solve_batch_learned(dg, qs[:128], model, args.frac, slots=args.slots)
        t = time.perf_counter(); lb, lst = solver.solve_batch_learned(dg, qs, model, args.frac, slots=args.slots)
        t_lb = time.perf_counter() - t
        t = time.perf_counter(); single = [solver.solve(dg, q) for q in qs[:64]]; t_s = time.perf_counter() - t
        for i in range(len(qs)):
            assert pb[i].cost == sb[i].cost == lb[i].cost, i
            if i < 64:
                assert single[i].cost == pb[i].cost, i
                assert gst.validate_tree(G, qs[i], lb[i].eu, lb[i].ev, lb[i].root) == lb[i].cost
        out.update({"metric": "exact-cost GST queries/s, 256 mixed queries (C4)", "value": round(len(qs) / t_lb, 3),
                    "unit": "queries/s", "dtype": "u32",
                    "config": {"workload": "BA V=200k m=5 int32 w[1,1000], 256 queries, g in 3..8, sizes 5..500",
                               "mode": f"learned-UB batched (top {args.frac:.0%}), {lst['slots']} slots"},
                    "learned_batched": {"value": round(len(qs) / t_lb, 3), "p50_ms": round(lst["latency_p50_ms"], 2),
                                        "p99_ms": round(lst["latency_p99_ms"], 2),
                                        "fallback_rate": round(sum(r.fell_back for r in lb) / len(lb), 4)},
                    "plain_batched": {"value": round(len(qs) / t_pb, 3), "p50_ms": round(pst["latency_p50_ms"], 2),
                                      "p99_ms": round(pst["latency_p99_ms"], 2)},
                    "spider_batched": {"value": round(len(qs) / t_sb, 3), "p50_ms": round(sst["latency_p50_ms"], 2),
                                       "p99_ms": round(sst["latency_p99_ms"], 2)},
                    "plain_one_at_a_time_first64": {"value": round(64 / t_s, 3)}})
    else:
        if args.workload == "rmat22":
            # float R-MAT scale 22 (a,b,c,d = .57,.19,.19,.05), edge factor 16, lognormal(0,1) weights, g=7
            G = gst.gen_rmat(22, 16, weights="lognormal", seed=22)
            g, smin, smax, nq, frac = 7, 20, 500, 3, 0.03
            dtype, metric = "f32", "exact-cost GST queries/s, float R-MAT scale 22 g=7 (C3)"
        else:
            # int R-MAT scale 21, weights [1,1000], g=8 groups of 10-300, learned bound on the top 5%
            G = gst.gen_rmat(21, 16, weights="int", wlo=1, whi=1000, seed=21)
            g, smin, smax, nq, frac = 8, 10, 300, 8, 0.05
            dtype, metric = "u32", "exact-cost GST queries/s, learned bound vs plain, int R-MAT scale 21 g=8 (C5)"
        log(f"graph V={G.n} arcs={G.m}")
        dg = solver.upload(G)
        log("uploaded")
        qs, k = [], 0
        while len(qs) < nq:   # R-MAT leaves isolated vertices: keep queries that have a tree
            q = gst.gen_groups(G.n, g, smin, smax, seed=7000 + k)
            k += 1
            try:
                solver.solve(dg, q, want_tree=False, auto_bound=True)
                qs.append(q)
            except gst.GstError:
                pass
        log(f"{nq} feasible queries ({k} drawn)")
        t = time.perf_counter(); plain = [solver.solve(dg, q) for q in qs]; t_p = time.perf_counter() - t
        log(f"plain {nq / t_p:.3f} q/s")
        t = time.perf_counter(); spider = [solver.solve(dg, q, auto_bound=True) for q in qs]; t_s = time.perf_counter() - t
        t = time.perf_counter(); lone = [solver.solve_learned(dg, q, model, frac) for q in qs]; t_l1 = time.perf_counter() - t
        # one slot per query (a query has one job in flight at a time); the engine's slots, captured
        # steps and subgraph buffers are set up by a warm-up call outside the timed region
        solver.solve_batch_learned(dg, qs, model, frac, slots=nq)
        t = time.perf_counter(); lb, lst = solver.solve_batch_learned(dg, qs, model, frac, slots=nq)
        t_lb = time.perf_counter() - t
        log(f"learned single {nq / t_l1:.3f} q/s, batched {nq / t_lb:.3f} q/s")
        rel = 1e-6 if dtype == "f32" else 0.0
        for i, q in enumerate(qs):
            c = plain[i].cost
            for r in (spider[i], lone[i], lb[i]):
                assert abs(r.cost - c) <= rel * c, (i, r.cost, c)
            tc = gst.validate_tree(G, q, plain[i].eu, plain[i].ev, plain[i].root)
            assert abs(tc - c) <= 1e-6 * max(c, 1.0), (i, tc, c)
        gms, mms = _rates(plain, "t_grow_ms"), _rates(plain, "t_merge_ms")
        gb, mb = sum(r.stats["grow_bytes"] for r in plain), sum(r.stats["merge_bytes"] for r in plain)
        gl = sum(r.stats["grow_launches"] for r in plain)
        ratios = [r.stats["ub"] / p.cost for r, p in zip(lone, plain) if r.stats["ub"] > 0]
        out.update({"metric": metric, "value": round(nq / t_lb, 4), "unit": "queries/s", "dtype": dtype,
                    "config": {"workload": f"R-MAT V={G.n} arcs={G.m} g={g} sizes {smin}-{smax}, {nq} queries",
                               "mode": f"learned-UB batched (top {frac:.0%}), {nq} slots"},
                    "plain_dpbf": {"value": round(nq / t_p, 4), "ms_per_query": round(t_p / nq * 1e3, 2)},
                    "spider_bound": {"value": round(nq / t_s, 4)},
                    "learned_one_at_a_time": {"value": round(nq / t_l1, 4)},
                    "learned": {"fallback_rate": round(sum(r.fell_back for r in lb) / nq, 4),
                                "ub_over_opt_mean": round(float(np.mean(ratios)), 5) if ratios else None},
                    "speedup_vs_plain": round(t_p / t_lb, 3),
                    "roofline": {"bound": "hbm", "kernel": "k_grow_round (plain)", "achieved": round(gbs(gb, gms), 2),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs(gb, gms) / HBM_PEAK_GBS, 5),
                                 "launches": gl, "ms_per_launch": round(gms / max(gl, 1), 5)},
                    "merge_roofline": {"bound": "hbm", "kernel": "k_merge (plain)", "achieved": round(gbs(mb, mms), 2),
                                       "frac": round(gbs(mb, mms) / HBM_PEAK_GBS, 5)},
                    "state_relax_per_s": round(sum(r.stats["grow_relax"] for r in plain) / (gms * 1e-3), 1)})
    print(json.dumps(out), flush=True)
    model.free()
    solver.close()


if __name__ == "__main__":
    main()
