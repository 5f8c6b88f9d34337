"""Training and evaluation of the learned upper bound (network kernels: csrc/gst_mpnn.hip).

Training data are exact solutions of the GPU DPBF itself: for every query on a fresh seeded
Barabasi-Albert graph the label of a vertex is 1 when it lies on the returned optimal tree.

    python -m gpu4gst_code_amd.learned train --queries 20000 --out gpu4gst_code_amd/models/mpnn_ba50k.npy
    python -m gpu4gst_code_amd.learned eval --model gpu4gst_code_amd/models/mpnn_ba50k.npy
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import gpu4gst_code_amd as gst  # noqa: E402

DEFAULT_MODEL = os.path.join(HERE, "models", "mpnn_ba50k.npy")


def tree_labels(n: int, r: "gst.Result") -> np.ndarray:
    y = np.zeros(n, np.uint8)
    if len(r.eu):
        y[r.eu] = 1
        y[r.ev] = 1
    if r.root >= 0:
        y[r.root] = 1
    return y


def query_stream(seed: int, n_graphs: int, per_graph: int, n: int, g_lo: int, g_hi: int, s_lo: int, s_hi: int):
    """(graph index, group set) pairs; graph i is BA(n, 5, w in [1,1000], seed + i)."""
    rng = np.random.default_rng(seed)
    for gi in range(n_graphs):
        for q in range(per_graph):
            g = int(rng.integers(g_lo, g_hi + 1))
            yield gi, g, int(rng.integers(1 << 62))


def train(solver, queries: int = 20000, n: int = 50000, n_graphs: int = 40, epochs: int = 3, lr: float = 2e-3,
          batch: int = 16, pos_weight: float = 300.0, seed: int = 7, log=print, model=None):
    per_graph = (queries + n_graphs - 1) // n_graphs
    model = model or gst.Model(solver, seed=seed)
    data = []   # (graph index, groups, labels)
    graphs = []
    t0 = time.time()
    solve_ms = 0.0
    for gi in range(n_graphs):
        G = gst.gen_barabasi_albert(n, 5, 1, 1000, seed=seed * 1000 + gi)
        dg = solver.upload(G)
        graphs.append(dg)
        rng = np.random.default_rng(seed * 7919 + gi)
        for q in range(per_graph):
            if len(data) >= queries:
                break
            g = int(rng.integers(3, 9))
            groups = gst.gen_groups(n, g, 10, 300, seed=int(rng.integers(1 << 62)))
            t = time.time()
            r = solver.solve(dg, groups)
            solve_ms += (time.time() - t) * 1e3
            data.append((gi, groups, tree_labels(n, r)))
        log(f"labels: graph {gi + 1}/{n_graphs}, {len(data)} queries, {time.time() - t0:.1f} s "
            f"(GPU DPBF {solve_ms / max(1, len(data)):.2f} ms/query)")
    order = np.arange(len(data))
    rng = np.random.default_rng(seed)
    history = []
    for ep in range(epochs):
        rng.shuffle(order)
        tot, cnt = 0.0, 0
        lr_ep = lr * (0.5 ** ep)
        for bi, i in enumerate(order):
            gi, groups, y = data[i]
            loss, _ = model.loss_grad(graphs[gi], groups, y, pos_weight)
            tot += loss
            cnt += 1
            if (bi + 1) % batch == 0:
                model.adam_step(lr_ep)
        model.adam_step(lr_ep)
        history.append(tot / cnt)
        log(f"epoch {ep + 1}/{epochs}: mean loss {tot / cnt:.5f} ({time.time() - t0:.1f} s)")
    return model, {"queries": len(data), "graphs": n_graphs, "n": n, "epochs": epochs, "loss": history,
                   "label_ms_per_query": solve_ms / max(1, len(data)), "seconds": time.time() - t0}


def recall_at(scores: np.ndarray, labels: np.ndarray, frac: float) -> float:
    k = max(1, int(np.ceil(frac * scores.shape[0])))
    top = np.argpartition(-scores, k - 1)[:k]
    pos = labels.sum()
    return float(labels[top].sum() / pos) if pos else 1.0


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--queries", type=int, default=20000)
    t.add_argument("--graphs", type=int, default=40)
    t.add_argument("--n", type=int, default=50000)
    t.add_argument("--epochs", type=int, default=3)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--pos-weight", type=float, default=300.0)
    t.add_argument("--out", default=DEFAULT_MODEL)
    args = ap.parse_args()
    if args.cmd == "train":
        with gst.Solver(0) as s:
            model, info = train(s, args.queries, args.n, args.graphs, args.epochs, args.lr, pos_weight=args.pos_weight,
                                log=lambda m: print(m, flush=True))
            os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
            model.save(args.out)
            with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
                json.dump(info, f, indent=1)
            print(json.dumps(info), flush=True)


if __name__ == "__main__":
    main()
