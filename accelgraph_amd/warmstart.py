"""Learned warm start: data collection, training and evaluation drivers.

    python -m accelgraph_amd.warmstart train --graphs 200 --out weights.json
    python -m accelgraph_amd.warmstart eval --weights weights.json --scale 22

Training data: seeded RMAT (a=0.57 b=0.19 c=0.19) and Erdos-Renyi graphs, alternating, scale
uniform over the given range, edge factor 16. For each graph the GPU computes the five features
(ag_mlp_features) and the plain solver's converged scaled ranks (L1 <= 1e-9 from the uniform
start); the target is their log ratio to the scaled ranks after two iterations. The network is
trained on the GPU (ag_mlp_train).
Evaluation compares time and iterations to L1 <= 1e-6 from the uniform start against the
warm start (features + inference + the fallback check included in its time).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import DEFAULT_SEED, INIT_UNIFORM, MLP, STOP_L1, Device, Graph


def training_graph(i: int, scale: int, seed_base: int = DEFAULT_SEED) -> Graph:
    if i % 2 == 0:
        return Graph.rmat(scale, seed=seed_base + i)
    n = 1 << scale
    return Graph.erdos_renyi(n, 16 * n, seed=seed_base + i)


def collect(dev: Device, g: Graph, max_samples: int = 0, rng=None):
    """Features of one graph's vertices and the targets log(r*) - log(r2): converged scaled
    rank over the scaled rank after two iterations (feature 4 is log(r2))."""
    dg = dev.upload(g)
    X = dg.features()
    run = dg.pagerank(max_iterations=400, stop_rule=STOP_L1, l1_tolerance=1e-9,
                      init=INIT_UNIFORM, history=False, scaled=True)
    y = (np.log(run.scaled.astype(np.float64)) - X[:, 4].astype(np.float64)).astype(np.float32)
    dg.close()
    if max_samples and g.n > max_samples:
        rng = rng or np.random.default_rng(0)
        idx = rng.choice(g.n, max_samples, replace=False)
        X, y = X[idx], y[idx]
    return X, y


def train(dev: Device, count: int = 200, lo: int = 14, hi: int = 18, epochs: int = 8,
          batch: int = 16384, lr: float = 2e-3, max_per_graph: int = 100000, seed: int = 0,
          log=print):
    rng = np.random.default_rng(seed)
    scales = rng.integers(lo, hi + 1, count)
    Xs, ys = [], []
    t0 = time.perf_counter()
    for i, s in enumerate(scales):
        g = training_graph(i, int(s))
        X, y = collect(dev, g, max_per_graph, rng)
        Xs.append(X)
        ys.append(y)
        g.close()
    X = np.concatenate(Xs)
    y = np.concatenate(ys)
    log(f"collected {X.shape[0]} samples from {count} graphs in {time.perf_counter() - t0:.1f} s")
    m = MLP(dev, seed=seed)
    m.fit_normalization(X, y)
    t0 = time.perf_counter()
    losses = m.train(X, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    log(f"trained {epochs} epochs in {time.perf_counter() - t0:.1f} s; loss per epoch "
        f"{[round(float(v), 5) for v in losses]}")
    return m, losses, (X.shape[0], count)


def evaluate(dev: Device, g: Graph, m: MLP, f64: np.ndarray | None = None) -> dict:
    """Time and iterations to L1 <= 1e-6: uniform start versus the learned warm start."""
    dg = dev.upload(g)
    kw = dict(max_iterations=400, stop_rule=STOP_L1, l1_tolerance=1e-6, init=INIT_UNIFORM,
              history=True)
    dg.pagerank(**kw)                                   # warm the caches / code paths
    plain = dg.pagerank(**kw)
    warm = dg.pagerank(warm=m, **kw)
    dg.close()
    plain_ms = float(plain.hist_ms.sum())
    warm_ms = float(warm.warm.ms + warm.hist_ms[3:].sum())
    out = {"n": g.n, "m": g.m,
           "uniform": {"iterations": plain.executed, "ms": round(plain_ms, 3)},
           "warm": {"iterations": warm.executed, "ms": round(warm_ms, 3),
                    "fallback": int(warm.warm.fallback),
                    "uniform_l1_after_2": warm.warm.uniform_l1,
                    "warm_l1_after_1": warm.warm.warm_l1},
           "speedup": round(plain_ms / max(warm_ms, 1e-9), 3)}
    if f64 is not None:
        out["uniform"]["max_abs_err_vs_f64"] = float(np.abs(plain.ranks - f64).max())
        out["warm"]["max_abs_err_vs_f64"] = float(np.abs(warm.ranks - f64).max())
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--graphs", type=int, default=200)
    t.add_argument("--lo", type=int, default=14)
    t.add_argument("--hi", type=int, default=18)
    t.add_argument("--epochs", type=int, default=8)
    t.add_argument("--batch", type=int, default=16384)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--out", required=True)
    e = sub.add_parser("eval")
    e.add_argument("--weights", required=True)
    e.add_argument("--scale", type=int, action="append")
    a = ap.parse_args(argv)
    dev = Device(0)
    if a.cmd == "train":
        m, losses, (S, G) = train(dev, a.graphs, a.lo, a.hi, a.epochs, a.batch, a.lr)
        m.save(a.out)
        with open(a.out) as f:
            d = json.load(f)
        d["training"] = {"graphs": G, "samples": S, "scales": f"{a.lo}-{a.hi}",
                         "epochs": a.epochs, "loss_per_epoch": [float(v) for v in losses]}
        with open(a.out, "w") as f:
            json.dump(d, f)
    else:
        m = MLP(dev)
        m.load(a.weights)
        for s in a.scale or [22]:
            g = Graph.rmat(s)
            print(json.dumps({"scale": s, **evaluate(dev, g, m)}), flush=True)
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
