"""Exact group Steiner tree queries/s on one MI355X.

Workload (one step = one exact GST query, solved through the C-ABI with inputs resident in HBM):
  Barabasi-Albert V=200,000, m=5 (~1.0M undirected edges), int32 weights uniform in [1,1000],
  g=6 groups of sizes uniform in [10,200]; query q uses group seed 1000+q (64 distinct queries,
  cycled).  Plain DPBF (no bound) is the headline `value`; the same queries are also reported with
  the learned upper bound when a trained model file is present.

Printed: one JSON line with the metric, the grow-kernel roofline (algorithmic bytes / kernel time
vs 8 TB/s), the merge-kernel roofline, state relaxations/s, and a CPU baseline (the oracle's
label-setting DPBF, run on the host cores over a bounded sample of the same queries).
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
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--n", type=int, default=200_000)
    ap.add_argument("--groups", type=int, default=6)
    ap.add_argument("--cpu-sample", type=int, default=16, help="queries in the CPU-baseline sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="MiB copied by one plain device copy before the run (counter calibration)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent queries would shard one-per-GPU with no communication)", file=sys.stderr)
        sys.exit(2)
    import gpu4gst_code_amd as gst

    G = gst.gen_barabasi_albert(args.n, 5, 1, 1000, seed=1)
    queries = [gst.gen_groups(G.n, args.groups, 10, 200, seed=1000 + q) for q in range(64)]
    solver = gst.Solver(0)
    dg = solver.upload(G)

    if args.calibrate_copy > 0:
        import ctypes
        calib = ctypes.CDLL(os.path.join(ROOT, "gpu4gst_code_amd", "libgst_calib.so"))
        calib.gst_calib_copy.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int]
        assert calib.gst_calib_copy(0, args.calibrate_copy << 20, 1) == 0

    for i in range(args.warmup):
        solver.solve(dg, queries[i % 64])

    # every solve ends with a stream synchronisation, so the host clock brackets device work
    t0 = time.perf_counter()
    results = []
    for i in range(args.steps):
        results.append(solver.solve(dg, queries[i % 64]))
    wall = time.perf_counter() - t0

    qps = args.steps / wall
    grow_ms = sum(r.stats["t_grow_ms"] for r in results)
    merge_ms = sum(r.stats["t_merge_ms"] for r in results)
    grow_bytes = sum(r.stats["grow_bytes"] for r in results)
    merge_bytes = sum(r.stats["merge_bytes"] for r in results)
    relax = sum(r.stats["grow_relax"] for r in results)
    grow_launches = sum(r.stats["grow_launches"] for r in results)
    grow_gbs = grow_bytes / (grow_ms * 1e-3) / 1e9 if grow_ms > 0 else 0.0
    merge_gbs = merge_bytes / (merge_ms * 1e-3) / 1e9 if merge_ms > 0 else 0.0

    # every tree is checked (cheap, host side)
    for i, r in enumerate(results[: min(len(results), 64)]):
        tc = gst.validate_tree(G, queries[i % 64], r.eu, r.ev, r.root)
        assert tc == r.cost, (i, tc, r.cost)

    cpu = None
    if args.cpu_sample > 0:
        import oracle
        sample = queries[: args.cpu_sample]
        nth = min(args.cpu_threads, len(sample), os.cpu_count() or 1)
        t1 = time.perf_counter()
        costs = oracle.dpbf_batch(G, sample, nthreads=nth)
        cw = time.perf_counter() - t1
        for i, c in enumerate(costs):
            assert c == results[i % len(results)].cost or i >= len(results), (i, c)
        cpu = {"value": round(len(sample) / cw, 4), "unit": "queries/s", "cores": nth, "kind": "port",
               "sample": f"{len(sample)} of the 64 queries, label-setting DPBF (oracle/dpbf_oracle.c), "
                         f"{nth} threads, {cw:.1f} s wall"}

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "dpbf_plain_ba200k_g6_traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath))["traffic_bytes_per_dispatch"]

    line = {
        "metric": "exact-cost GST queries/s (plain DPBF, BA V=200k g=6)",
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
                               "64 queries one at a time", "mode": "plain DPBF"},
        "state_relax_per_s": round(relax / (grow_ms * 1e-3), 1) if grow_ms > 0 else None,
        "roofline": {"bound": "hbm", "achieved": round(grow_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(grow_gbs / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "kernel": "k_grow_round", "launches": grow_launches,
                     "bytes_per_launch": round(grow_bytes / max(grow_launches, 1), 1),
                     "ms_per_launch": round(grow_ms / max(grow_launches, 1), 5)},
        "merge_roofline": {"bound": "hbm", "achieved": round(merge_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(merge_gbs / HBM_PEAK_GBS, 5), "kernel": "k_merge"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    solver.close()


if __name__ == "__main__":
    main()
