"""PageRank-pull benchmark on one MI355X.

Workload (default): PageRank pull, FP32 ranks, on a synthetic directed RMAT graph of scale 25
(33,554,432 vertices, edge factor 16, a=0.57 b=0.19 c=0.19 d=0.05, seed 27491095, self-loops
and duplicate edges removed). One step = one PageRank iteration over the whole graph (gather
over the inverse CSR, damping update, L1/active reduction, next contribution), with the graph
already resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 25] [--graph rmat|er]

Prints one JSON line: value = GTEPS (edges / second / 1e9), the HBM roofline of the
iteration kernel (algorithmic bytes 8M + 20N + 4 per launch over the kernel's average
duration, measured with HIP events on the library's stream), and the CPU baseline (the
oracle's OpenMP restatement of the reference pull, timed on this host's cores on a bounded
number of iterations of the same graph).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# host cores for the CPU legs, read before any OpenMP runtime loads: with OMP_PROC_BIND set,
# the runtime pins the main thread to one core and the affinity mask read later shrinks to it
HOST_THREADS = len(os.sched_getaffinity(0))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


HBM_PEAK_GBS = 8000.0
METRIC = "PageRank-pull GTEPS + HBM GB/s; time-to-L1≤1e-6 plain vs warm-start/batched"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _tmpdir() -> str:
    import tempfile
    return tempfile.mkdtemp(prefix="ag_bench_")


def algorithmic_bytes(n: int, m: int, kind: str) -> int:
    # inv offsets 4(N+1), source ids 4M, one contribution per edge, out-degree 4N,
    # rank read 4N, rank write 4N, contribution write 4N
    if kind in ("bf16", "f16"):
        return 6 * m + 18 * n + 4
    if kind in ("q32", "q16", "q8"):
        # B-bit codes gathered per edge, f32 contributions written by the update, and the
        # code pass: f32 contributions read, codes written
        c = int(kind[1:]) // 8
        return (4 + c) * m + 20 * n + 4 + (4 + c) * n
    return 8 * m + 20 * n + 4


def int_kind_iteration(ag, dev, dg, m: int, kind: int, nbytes: int) -> dict:
    """Kernel time of one full iteration of an integer-sum kind (no early stop)."""
    s = dg.session(kind=kind, max_iterations=3 + 10, stop_rule=ag.STOP_NONE,
                   init=ag.INIT_UNIFORM)
    s.step(3)
    dev.sync()
    s.step(10)
    dev.sync()
    ms, _ = s.last_step_kernel_ms()
    s.end()
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"ms_per_iteration": round(ms, 4), "gteps": round(m / (ms * 1e-3) / 1e9, 2),
            "algorithmic_bytes": nbytes, "gb_per_s": round(gbs, 1),
            "roofline_frac": round(gbs / HBM_PEAK_GBS, 4)}


def cpu_rows(g, quant: dict, directions: dict) -> None:
    """The oracle's OpenMP restatements of the same rows on the host cores, on bounded samples
    (2 iterations per quantised kind, 1 per push kind, the first 5 data-driven iterations),
    written beside each GPU figure."""
    import oracle
    threads = HOST_THREADS
    inv = (g.out_degree, g.inv_idx, g.in_degree, g.inv_src)
    fwd = (g.edges_idx, g.out_degree, g.fwd_dest)
    for b, kind in ((32, oracle.KIND_QUANT32), (16, oracle.KIND_QUANT16), (8, oracle.KIND_QUANT8)):
        t0 = time.perf_counter()
        oracle.pagerank_pull(*inv, kind=kind, max_iterations=2, stop_rule=2, threads=threads)
        ms = (time.perf_counter() - t0) * 1e3 / 2
        row = quant[f"quant{b}"]
        row["cpu_ms_per_iteration"] = round(ms, 2)
        row["gpu_speedup"] = round(ms / row["ms_per_iteration"], 1)
    for name, kind in (("push_f32", 0), ("push_fixed64", 64)):
        t0 = time.perf_counter()
        oracle.pagerank_push(g.out_degree, *fwd, kind=kind, max_iterations=1, stop_rule=2,
                             threads=threads)
        ms = (time.perf_counter() - t0) * 1e3
        row = directions[name]
        row["cpu_ms_per_iteration"] = round(ms, 2)
        row["gpu_speedup"] = round(ms / row["ms_per_iteration"], 1)
    t0 = time.perf_counter()
    oracle.pagerank_dd_pull(*inv, *fwd, max_iterations=5, epsilon=1e-4, threads=threads)
    ms = (time.perf_counter() - t0) * 1e3
    row = directions["data_driven_pull_first_5_iterations"]
    row["cpu_ms"] = round(ms, 1)
    row["gpu_speedup"] = round(ms / row["loop_ms"], 1)
    directions["cpu_cores"] = threads


def directions_section(ag, dev, dg, m: int, n: int) -> dict:
    """Push iterations (propagation-blocked, FP32 and fixed-64) and the data-driven pull
    against the plain pull under the reference's stop rule (epsilon 1e-4, scaled units)."""
    out = {}
    t0 = time.perf_counter()
    s = dg.session(direction=ag.DIR_PUSH, max_iterations=1, stop_rule=ag.STOP_NONE)
    dev.sync()
    s.end()
    out["forward_csr_and_push_bins_build_s"] = round(time.perf_counter() - t0, 3)
    for name, kind in (("push_f32", ag.KIND_F32), ("push_fixed64", ag.KIND_FIXED64)):
        s = dg.session(kind=kind, direction=ag.DIR_PUSH, max_iterations=13,
                       stop_rule=ag.STOP_NONE, init=ag.INIT_UNIFORM)
        s.step(3)
        dev.sync()
        s.step(10)
        dev.sync()
        ms, _ = s.last_step_kernel_ms()
        s.end()
        # phase 1: source id + slot per out-edge, contributions read once, one value written
        # per slot; phase 2: 16-bit row + value per slot, then the update of every row (rank
        # read + written, out-degree, next contribution)
        cb = 4 if kind == ag.KIND_F32 else 8
        nbytes = (10 + 2 * cb) * m + (12 + 2 * cb) * n
        out[name] = {"ms_per_iteration": round(ms, 4), "gteps": round(m / (ms * 1e-3) / 1e9, 2),
                     "algorithmic_bytes": nbytes,
                     "gb_per_s": round(nbytes / (ms * 1e-3) / 1e9, 1)}
    for name, d in (("pull", ag.DIR_PULL), ("data_driven_pull", ag.DIR_DATA_DRIVEN)):
        run = dg.pagerank(direction=d, max_iterations=200, epsilon=1e-4)
        out[f"{name}_to_reference_stop"] = {"iterations": run.iterations,
                                            "loop_ms": round(run.loop_ms, 2),
                                            "error_total": run.error_total}
    run = dg.pagerank(direction=ag.DIR_DATA_DRIVEN, max_iterations=5, epsilon=1e-4,
                      stop_rule=ag.STOP_NONE)
    out["data_driven_pull_first_5_iterations"] = {"loop_ms": round(run.loop_ms, 2)}
    return out


def cpu_baseline(g, budget_s: float) -> dict:
    import oracle
    threads = HOST_THREADS
    args = (g.out_degree, g.inv_idx, g.in_degree, g.inv_src)
    t0 = time.perf_counter()
    oracle.pagerank_pull(*args, max_iterations=1, stop_rule=2, threads=threads)
    t1 = time.perf_counter() - t0
    k = int(max(1, min(20, budget_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    oracle.pagerank_pull(*args, max_iterations=k, stop_rule=2, threads=threads)
    dt = (time.perf_counter() - t0) / k
    return {"value": round(g.m / dt / 1e9, 4), "unit": "GTEPS", "cores": threads,
            "cpu_model": cpu_model(), "kind": "port",
            "sample": f"{k} iterations of the OpenMP restatement (schedule(dynamic,1024)) of "
                      f"the reference pull on the same graph, {dt * 1e3:.1f} ms/iteration",
            "ms_per_iteration": round(dt * 1e3, 3)}


def time_to_l1(ag, dg, g, f64) -> dict:
    """Plain FP32 vs BF16-stored / FP32-refined time to L1 <= 1e-6 from the uniform start,
    each checked against the FP64 oracle."""
    import numpy as np
    kw = dict(max_iterations=300, stop_rule=ag.STOP_L1, l1_tolerance=1e-6,
              init=ag.INIT_UNIFORM, history=True)
    plain = dg.pagerank(**kw)
    mixed = dg.pagerank(storage=ag.STORE_BF16, refine_l1=1e-3, **kw)
    out = {"plain_f32": {"ms": round(float(plain.hist_ms.sum()), 3),
                         "iterations": plain.executed, "final_l1": plain.error_total},
           "bf16_refined": {"ms": round(float(mixed.hist_ms.sum()), 3),
                            "iterations": mixed.executed, "final_l1": mixed.error_total,
                            "refine_iteration": mixed.refine_iteration}}
    out["bf16_speedup"] = round(out["plain_f32"]["ms"] / max(out["bf16_refined"]["ms"], 1e-9), 3)
    if f64 is not None:
        for key, run in (("plain_f32", plain), ("bf16_refined", mixed)):
            d = np.abs(run.ranks.astype(np.float64) - f64)
            out[key]["max_abs_err_vs_f64"] = float(d.max())
            out[key]["l1_err_vs_f64"] = float(d.sum())
    return out


def warm_section(ag, dev, g, f64, cpu: bool) -> dict:
    """Learned warm start on the bench graph and on the held-out RMAT-22 and RMAT-24 graphs
    (the network was trained on graphs of scale 14-18), time-to-L1 <= 1e-6 including
    features and inference, vs the uniform start, each with its error against the FP64
    oracle (CPU leg)."""
    from accelgraph_amd import warmstart
    wpath = os.path.join(ROOT, "accelgraph_amd", "warmstart_weights.json")
    m = ag.MLP(dev)
    m.load(wpath)
    res = warmstart.evaluate(dev, g, m, f64)
    for scale in (22, 24):
        h = ag.Graph.rmat_gpu(dev, scale)
        hf = None
        if cpu:
            import oracle
            hf = oracle.pagerank_f64(h.out_degree, h.inv_idx, h.in_degree, h.inv_src,
                                     tol=1e-13, max_iterations=150, threads=HOST_THREADS)
        res[f"held_out_rmat{scale}"] = warmstart.evaluate(dev, h, m, hf)
        h.close()
    with open(wpath) as f:
        res["training"] = json.load(f).get("training")
    m.close()
    return res


def batched_section(ag, dev, count: int, lo: int, hi: int, cpu: bool) -> dict:
    """count RMAT graphs with vertex counts 2^s log-uniform, i.e. scale s uniform over the
    integers [lo, hi] (seeded), to L1 <= 1e-6: one persistent hipGraph launch versus the plain
    single-graph GPU path run graph by graph; each graph's distance to its FP64 solution
    (bounded through the FP64 residual, CPU leg)."""
    import numpy as np
    rng = np.random.default_rng(ag.DEFAULT_SEED)
    scales = rng.integers(lo, hi + 1, count)
    t0 = time.perf_counter()
    graphs = [ag.Graph.rmat(int(s), seed=ag.DEFAULT_SEED + i) for i, s in enumerate(scales)]
    log(f"batch: built {count} graphs (scales {lo}-{hi}, {sum(g.m for g in graphs)} edges) in "
        f"{time.perf_counter() - t0:.1f} s")
    b = ag.Batch(dev, graphs)
    kw = dict(max_iterations=300, stop_rule=ag.STOP_L1, l1_tolerance=1e-6, init=ag.INIT_UNIFORM)
    # part size and heavy-row threshold picked by the autotuner (timed on this solve)
    from accelgraph_amd import autotune
    tuned, rep = autotune.autotune_batch(b, repeats=2, passes=1, force=True,
                                         table_path=os.path.join(_tmpdir(), "tuning.json"))
    b.pagerank(use_hip_graph=1, **kw)                       # capture + warm (defaults)
    _, _, _, default_ms = b.pagerank(use_hip_graph=1, **kw)
    b.pagerank(use_hip_graph=1, **tuned, **kw)              # capture + warm (tuned)
    ranks, its, err, ms = b.pagerank(use_hip_graph=1, **tuned, **kw)
    dgs = [dev.upload(g) for g in graphs]
    for dg in dgs[:4]:
        dg.pagerank(history=False, **kw)                    # warm
    dev.sync()
    t0 = time.perf_counter()
    singles = [dg.pagerank(history=False, **kw) for dg in dgs]
    plain_ms = (time.perf_counter() - t0) * 1e3
    worst = max(float(np.abs(s.ranks - r).max()) for s, r in zip(singles, ranks))
    for dg in dgs:
        dg.close()
    b.close()
    out = {"graphs": count, "scales": f"{lo}-{hi} (2^s log-uniform)", "edges": int(sum(g.m for g in graphs)),
           "batched_ms": round(ms, 3), "plain_loop_ms": round(plain_ms, 3),
           "autotuned": {"options": tuned, "ms": round(ms, 3),
                         "default_options": rep["default"], "default_ms": round(default_ms, 3),
                         "trials": rep["trials"]},
           "speedup": round(plain_ms / max(ms, 1e-9), 2),
           "max_iterations": int(its.max()), "max_final_l1": float(err.max()),
           "max_abs_diff_vs_single_graph_path": worst}
    if cpu:
        # per-graph L1 distance to the FP64 solution (normalised units), bounded by the FP64
        # residual: ||r - r*||_1 <= ||p + 0.85 A r - r||_1 / 0.15
        import oracle
        rel = []
        for g, r in zip(graphs, ranks):
            res, nrm = oracle.ppr_residual(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                                           [ag.PPR_UNIFORM], r, threads=HOST_THREADS)
            rel.append(float(res[0] / 0.15))
        out["max_l1_err_vs_f64_bound"] = max(rel)
        out["median_l1_err_vs_f64_bound"] = float(np.median(rel))
    return out


def ppr_section(ag, dev, scale: int, k: int, cpu: bool) -> dict:
    """k-column personalised PageRank (column 0 uniform = plain PageRank, the others seeded
    uniformly at random by seed) to L1 <= 1e-6 in BF16, FP16 and FP32 storage (16-bit storage
    refined in FP32), one captured hipGraph solve each, versus k plain single-vector runs
    (k x one plain run's time); every column's distance to its FP64 solution is bounded
    through the FP64 residual (CPU leg)."""
    import numpy as np
    g = ag.Graph.rmat(scale)
    dg = dev.upload(g)
    rng = np.random.default_rng(ag.DEFAULT_SEED)
    seeds = np.concatenate([[ag.PPR_UNIFORM], rng.choice(g.n, k - 1, replace=False)])
    res = {}
    from accelgraph_amd import autotune
    tuned, rep = autotune.autotune_ppr(dg, seeds, iterations=6, repeats=2, force=True,
                                       table_path=os.path.join(_tmpdir(), "tuning.json"))
    res["autotuned"] = {"options": tuned, "bf16_ms_per_iteration": round(rep["ms_per_iteration"], 4),
                        "default_options": rep["default"],
                        "default_bf16_ms_per_iteration": round(rep["default_ms_per_iteration"], 4),
                        "trials": rep["trials"]}
    for name, st in (("bf16", ag.STORE_BF16), ("f16", ag.STORE_F16), ("f32", ag.STORE_F32)):
        dg.ppr(seeds, max_iterations=300, stop_rule=ag.STOP_L1, l1_tolerance=1e-6, storage=st,
               refine_l1=1e-3, **tuned)                      # capture + warm
        ranks, colerr, r = dg.ppr(seeds, max_iterations=300, stop_rule=ag.STOP_L1,
                                  l1_tolerance=1e-6, storage=st, refine_l1=1e-3, **tuned)
        res[name] = {"ms": round(r.loop_ms, 3), "iterations": r.executed,
                     "ms_per_iteration": round(r.loop_ms / max(r.executed, 1), 4),
                     "max_col_l1": r.max_error, "refine_iteration": r.refine_iteration}
        if cpu:
            # every column's L1 distance to its FP64 solution, bounded by the FP64 residual
            import oracle
            rr, nrm = oracle.ppr_residual(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                                          seeds, ranks, threads=HOST_THREADS)
            res[name]["max_col_l1_err_vs_f64_bound"] = float((rr / 0.15).max())
        if name == "bf16":
            col0 = ranks[:, 0].copy()
    plain = dg.pagerank(max_iterations=300, stop_rule=ag.STOP_L1, l1_tolerance=1e-6,
                        init=ag.INIT_UNIFORM, history=True)
    plain_ms = float(plain.hist_ms.sum())
    res["plain_single_ms"] = round(plain_ms, 3)
    res["k_plain_runs_ms"] = round(k * plain_ms, 3)
    res["speedup_vs_k_plain"] = round(k * plain_ms / max(res["bf16"]["ms"], 1e-9), 2)
    res["col0_max_abs_diff_vs_plain"] = float(np.abs(col0 - plain.ranks).max())
    res["graph"] = f"RMAT scale {scale} (n={g.n}, m={g.m}), {k} columns"
    dg.close()
    return res


def spmv_section(ag, dev, scale: int, reps: int, cpu: bool, pkw: dict) -> dict:
    """SpMV pull on a weighted RMAT graph: w = (k+1) 2^-24 and x = k 2^-24 with k uniform on
    [0, 2^24) from a seeded stream; FP32 (binary32 products, FP32 sums) and the RTL's Q16.16
    fixed-point datapath (u64 sums). Algorithmic bytes 12M + 8N + 4 (FP32), 12M + 12N + 4
    (fixed). The CPU leg times the oracle's OpenMP restatement on the same inputs."""
    import numpy as np
    g = ag.Graph.rmat(scale)
    rng = np.random.default_rng(ag.DEFAULT_SEED)
    g.set_weights(((rng.integers(0, 1 << 24, g.m) + 1) * 2.0 ** -24).astype(np.float32))
    x = (rng.integers(0, 1 << 24, g.n) * 2.0 ** -24).astype(np.float32)
    dg = dev.upload(g, weights=True, params=ag.plan_params(**pkw) if pkw else None)
    info = dg.info()
    out = {"graph": f"RMAT scale {scale} (n={g.n}, m={g.m}), weighted",
           "src_block_bits": info.src_block_bits, "sorted_tiles": bool(info.sorted_tiles)}
    xq = dev.fixed_encode(32, x).astype(np.uint32)
    for name, fn, arg, nb in (("f32", dg.spmv, x, 12 * g.m + 8 * g.n + 4),
                              ("fixed32", dg.spmv_fixed, xq, 12 * g.m + 12 * g.n + 4)):
        fn(arg, reps=2)                                     # warm
        _, ms = fn(arg, reps=reps)
        gbs = nb / (ms * 1e-3) / 1e9
        out[name] = {"ms": round(ms, 4), "gteps": round(g.m / (ms * 1e-3) / 1e9, 2),
                     "algorithmic_bytes": nb, "gb_per_s": round(gbs, 1),
                     "roofline_frac": round(gbs / HBM_PEAK_GBS, 4)}
    # push forms (SPMV.c:752-848, 1017-1136), propagation-blocked; bytes: phase 1 streams
    # source id, slot and weight per out-edge, reads x once and writes one product per slot;
    # phase 2 reads a 16-bit row and the product per slot and updates y (read + write)
    dg.spmv_push(x, reps=1)                  # warm (builds the forward CSR and the push bins)
    xq64 = dev.fixed_encode(64, x)
    for name, fn, arg, nb in (("push_f32", dg.spmv_push, x, 22 * g.m + 12 * g.n),
                              ("push_fixed64", dg.spmv_push_fixed, xq64, 30 * g.m + 24 * g.n)):
        _, ms = fn(arg, reps=reps)
        out[name] = {"ms": round(ms, 4), "gteps": round(g.m / (ms * 1e-3) / 1e9, 2),
                     "algorithmic_bytes": nb, "gb_per_s": round(nb / (ms * 1e-3) / 1e9, 1)}
    y, _ = dg.spmv(x)
    # fixed-32 PageRank iteration on the same graph (Q16.16 contributions, exact u64 sums)
    fs = dg.session(kind=ag.KIND_FIXED32, max_iterations=3 + 20, stop_rule=ag.STOP_NONE,
                    init=ag.INIT_UNIFORM)
    fs.step(3)
    dev.sync()
    fs.step(20)
    dev.sync()
    fms, _ = fs.last_step_kernel_ms()
    fs.end()
    fgbs = algorithmic_bytes(g.n, g.m, "f32") / (fms * 1e-3) / 1e9
    out["pagerank_fixed32_iteration"] = {"ms": round(fms, 4),
                                         "gteps": round(g.m / (fms * 1e-3) / 1e9, 2),
                                         "gb_per_s": round(fgbs, 1),
                                         "roofline_frac": round(fgbs / HBM_PEAK_GBS, 4)}
    dg.close()
    if cpu:   # CPU leg: the oracle's OpenMP restatement, also the FP32 result check
        import oracle
        threads = HOST_THREADS
        t0 = time.perf_counter()
        ref = oracle.spmv_f32(g.inv_idx, g.in_degree, g.inv_src, g.inv_weight, x, threads)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"ms": round(dt * 1e3, 2), "cores": threads, "kind": "port",
                               "gteps": round(g.m / dt / 1e9, 3)}
        out["f32"]["max_rel_err_vs_cpu"] = float(np.max(np.abs(y - ref) / np.maximum(np.abs(ref), 1e-30)))
        t0 = time.perf_counter()
        oracle.spmv_push_f32(g.edges_idx, g.out_degree, g.fwd_dest, g.fwd_weight, x, 1, threads)
        dt = time.perf_counter() - t0
        out["push_f32"]["cpu_ms"] = round(dt * 1e3, 2)
        out["push_f32"]["gpu_speedup"] = round(dt * 1e3 / out["push_f32"]["ms"], 1)
        # fixed-32 PageRank: 3 iterations of the restatement, per iteration
        t0 = time.perf_counter()
        oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src, kind=32,
                             max_iterations=3, stop_rule=2, init=1, threads=threads)
        dt = (time.perf_counter() - t0) / 3
        row = out["pagerank_fixed32_iteration"]
        row["cpu_ms_per_iteration"] = round(dt * 1e3, 2)
        row["gpu_speedup"] = round(dt * 1e3 / row["ms"], 1)
    return out


def config1_section(ag, dev, cpu: bool) -> dict:
    """RMAT-16, damping 0.85, uniform 1/N start, stop at L1 <= 1e-4 or 100 iterations: time to
    the stop (graph resident) and the error against the FP64 oracle (CPU leg)."""
    import numpy as np
    g = ag.Graph.rmat(16)
    dg = dev.upload(g)
    kw = dict(max_iterations=100, stop_rule=ag.STOP_L1, l1_tolerance=1e-4, init=ag.INIT_UNIFORM,
              history=True)
    dg.pagerank(**kw)                                    # warm the code paths
    run = dg.pagerank(**kw)
    dg.close()
    out = {"graph": f"RMAT scale 16 (n={g.n}, m={g.m})", "iterations": run.executed,
           "final_l1": run.error_total, "ms": round(float(run.hist_ms.sum()), 4),
           "ms_per_iteration": round(float(run.hist_ms.mean()), 4)}
    if cpu:
        import oracle
        threads = HOST_THREADS
        args = (g.out_degree, g.inv_idx, g.in_degree, g.inv_src)
        # FP64 after the same number of iterations from the same start: the FP32 rounding error
        f64 = oracle.pagerank_f64(*args, tol=0.0, max_iterations=run.executed, threads=threads)
        out["max_abs_err_vs_f64_same_iterations"] = float(np.abs(run.ranks - f64).max())
        # the same solve on the host cores: the oracle's OpenMP restatement of the reference
        # pull, same start and stop rule
        t0 = time.perf_counter()
        ref = oracle.pagerank_pull(*args, max_iterations=100, stop_rule=1, l1_tolerance=1e-4,
                                   init=1, threads=threads)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        out["cpu"] = {"ms": round(cpu_ms, 2), "iterations": int(ref.executed), "cores": threads,
                      "kind": "port", "gpu_speedup": round(cpu_ms / out["ms"], 1)}
    return out


def er_section(ag, dev, scale: int, steps: int, cpu: bool, pkw: dict, m_target: int) -> dict:
    """PageRank-pull FP32 iteration on an Erdos-Renyi graph with the bench graph's vertex and
    edge counts (uniform sources: no hot set, every gather is a cold random read), and the
    solve to L1 <= 1e-6 from the uniform start with its error against the FP64 oracle (CPU
    leg). The draws are the RMAT graph's edge count plus the expected self-loops and repeated
    pairs that simplification removes (draws / n + draws^2 / 2n^2, about m / 2^22 at
    scale 25), so both graphs of the configuration have the same N and, within a few dozen
    edges, the same M."""
    import numpy as np
    n = 1 << scale
    extra = int(m_target / n + m_target * m_target / (2.0 * n * n)) + 1
    g = ag.Graph.erdos_renyi_gpu(dev, n, m_target + extra)
    dg = dev.upload(g, params=ag.plan_params(**pkw) if pkw else None)
    sess = dg.session(max_iterations=steps + 3, stop_rule=ag.STOP_NONE, init=ag.INIT_UNIFORM)
    sess.step(3)
    dev.sync()
    sess.step(steps)
    dev.sync()
    ms, _ = sess.last_step_kernel_ms()
    sess.end()
    info = dg.info()
    kw = dict(max_iterations=300, stop_rule=ag.STOP_L1, l1_tolerance=1e-6,
              init=ag.INIT_UNIFORM, history=True)
    run = dg.pagerank(**kw)
    dg.close()
    nb = algorithmic_bytes(g.n, g.m, "f32")
    gbs = nb / (ms * 1e-3) / 1e9
    out = {"graph": f"Erdos-Renyi n=2^{scale}, {m_target + extra} draws, simplified "
                    f"(n={g.n}, m={g.m}; the RMAT graph's m={m_target})",
           "ms_per_iteration": round(ms, 4), "gteps": round(g.m / (ms * 1e-3) / 1e9, 2),
           "gb_per_s": round(gbs, 1), "roofline_frac": round(gbs / HBM_PEAK_GBS, 4),
           "src_block_bits": info.src_block_bits, "hot_segments": info.hot_segments,
           "time_to_l1_1e-6": {"ms": round(float(run.hist_ms.sum()), 3),
                               "iterations": run.executed, "final_l1": run.error_total}}
    if cpu:
        import oracle
        f64 = oracle.pagerank_f64(g.out_degree, g.inv_idx, g.in_degree, g.inv_src, tol=1e-13,
                                  max_iterations=150, threads=HOST_THREADS)
        d = np.abs(run.ranks.astype(np.float64) - f64)
        out["time_to_l1_1e-6"]["max_abs_err_vs_f64"] = float(d.max())
        out["time_to_l1_1e-6"]["l1_err_vs_f64"] = float(d.sum())
    g.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=int, default=25)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--graph", choices=["rmat", "er"], default="rmat")
    ap.add_argument("--storage", choices=["f32", "bf16", "f16"], default="f32")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ttl", action="store_true", help="skip the time-to-L1 runs")
    ap.add_argument("--f64-iters", type=int, default=150,
                    help="FP64 oracle iterations for the error report (0 = skip)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-graph batch")
    ap.add_argument("--no-warm", action="store_true", help="skip the learned warm start")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--no-ppr", action="store_true", help="skip the 64-column PPR")
    ap.add_argument("--ppr-scale", type=int, default=22)
    ap.add_argument("--no-spmv", action="store_true", help="skip the RMAT-22 SpMV")
    ap.add_argument("--spmv-scale", type=int, default=22)
    ap.add_argument("--no-er", action="store_true", help="skip the ER PageRank iteration")
    ap.add_argument("--hot-segments", type=int, default=-1,
                    help="plan parameter hot_segments for the bench graph (-1 = the default)")
    ap.add_argument("--tile-order", type=int, default=0,
                    help="plan parameter tile_order for the bench graphs (0 auto, 1 round-robin, "
                         "2 ticket)")
    ap.add_argument("--piece-nnz", type=int, default=0,
                    help="plan parameter piece_nnz for the bench graphs (0 = auto)")
    ap.add_argument("--src-block-bits", type=int, default=0,
                    help="plan parameter src_block_bits for the bench graphs (0 = auto)")
    ap.add_argument("--top-table", type=int, default=0,
                    help="plan parameter top_table for the bench graphs (0 auto, 1 off, 2 on)")
    ap.add_argument("--tile-sort", type=int, default=0,
                    help="plan parameter tile_sort for the bench graphs (0 auto, 1 off, 2 on)")
    ap.add_argument("--no-rows", action="store_true",
                    help="skip the fixed-32, quantised, push and data-driven rows")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (a sharded "
              "version would split destination rows across GPUs and all-gather the "
              "contribution vector every iteration). Run with --gpus 1.", file=sys.stderr)
        return 2

    import accelgraph_amd as ag

    t0 = time.perf_counter()
    if args.graph == "rmat":
        g = ag.Graph.rmat(args.scale, args.edge_factor)
        gdesc = (f"RMAT scale {args.scale}, edge factor {args.edge_factor}, a=0.57 b=0.19 "
                 f"c=0.19 d=0.05, seed {ag.DEFAULT_SEED}, simplified")
    else:
        n = 1 << args.scale
        g = ag.Graph.erdos_renyi(n, args.edge_factor * n)
        gdesc = (f"Erdos-Renyi n=2^{args.scale}, {args.edge_factor}*n draws, seed "
                 f"{ag.DEFAULT_SEED}, simplified")
    n, m = g.n, g.m
    log(f"graph built: n={n} m={m} in {time.perf_counter() - t0:.1f} s")

    cpu_build_s = time.perf_counter() - t0
    dev = ag.Device(0)
    copy_gbs = dev.copy_bandwidth(1 << 30, 10)
    # the same graph built on the GPU (generator + CSR + inverse CSR), checked array for array
    construction = None
    if args.graph == "rmat":
        import numpy as np
        t1 = time.perf_counter()
        gg = ag.Graph.rmat_gpu(dev, args.scale, args.edge_factor)
        gpu_build_s = time.perf_counter() - t1
        same = (gg.n, gg.m) == (n, m) and all(
            np.array_equal(getattr(gg, k), getattr(g, k))
            for k in ("out_degree", "edges_idx", "fwd_dest", "in_degree", "inv_idx", "inv_src"))
        gg.close()
        construction = {"host_s": round(cpu_build_s, 3), "gpu_s": round(gpu_build_s, 3),
                        "speedup": round(cpu_build_s / max(gpu_build_s, 1e-9), 2),
                        "identical_arrays": bool(same),
                        "host_cores": HOST_THREADS}
        log(f"graph construction: {construction}")
    log(f"device copy bandwidth {copy_gbs:.0f} GB/s (1 GiB, read + write)")
    t0 = time.perf_counter()
    pkw = {}
    if args.hot_segments >= 0:
        pkw["hot_segments"] = args.hot_segments
    if args.tile_order:
        pkw["tile_order"] = args.tile_order
    if args.piece_nnz:
        pkw["piece_nnz"] = args.piece_nnz
    if args.src_block_bits:
        pkw["src_block_bits"] = args.src_block_bits
    if args.top_table:
        pkw["top_table"] = args.top_table
    if args.tile_sort:
        pkw["tile_sort"] = args.tile_sort
    dg = dev.upload(g, params=ag.plan_params(**pkw) if pkw else None)
    info = dg.info()
    log(f"uploaded + planned in {time.perf_counter() - t0:.1f} s; bins rows="
        f"{list(info.bin_rows)} edges={list(info.bin_edges)} tiles={info.num_tiles} "
        f"segments={info.num_segments}")

    storage = {"f32": ag.STORE_F32, "bf16": ag.STORE_BF16, "f16": ag.STORE_F16}[args.storage]
    K, W = args.steps, args.warmup
    # fixed-iteration session: every launch is a full iteration (no early stop)
    sess = dg.session(max_iterations=W + K, stop_rule=ag.STOP_NONE, init=ag.INIT_UNIFORM,
                      storage=storage, refine_l1=0.0)
    if W:
        sess.step(W)
    dev.sync()
    t0 = time.perf_counter()
    sess.step(K)
    dev.sync()
    wall = time.perf_counter() - t0
    kern_ms, launches = sess.last_step_kernel_ms()
    sess.end()
    ms_per_step = wall * 1e3 / K
    gteps = m / (ms_per_step * 1e-3) / 1e9
    nbytes = algorithmic_bytes(n, m, args.storage)
    achieved = nbytes / (kern_ms * 1e-3) / 1e9
    log(f"{K} steps: {ms_per_step:.3f} ms/step wall, kernel avg {kern_ms:.3f} ms over "
        f"{launches} launches, {gteps:.2f} GTEPS, {achieved:.0f} GB/s algorithmic")

    fixed32 = quant = directions = None
    if not args.no_rows:
        # Q16.16 fixed-point iteration (the reference's pushpull 9 path: u32 contributions,
        # exact u64 sums), same graph, same algorithmic bytes
        fixed32 = int_kind_iteration(ag, dev, dg, m, ag.KIND_FIXED32,
                                     algorithmic_bytes(n, m, "f32"))
        log(f"fixed-32 iteration: {fixed32}")
        # quantised iterations (pushpull 4/12/13): B-bit codes gathered, plus the code pass
        quant = {f"quant{b}": int_kind_iteration(ag, dev, dg, m, kind,
                                                 algorithmic_bytes(n, m, f"q{b}"))
                 for b, kind in ((32, ag.KIND_QUANT32), (16, ag.KIND_QUANT16),
                                 (8, ag.KIND_QUANT8))}
        log(f"quantised iterations: {quant}")
        directions = directions_section(ag, dev, dg, m, n)
        log(f"push / data-driven: {directions}")

    # HBM traffic per launch is not measured in this run (counters need rocprofv3 runs of
    # their own): it is the figure of the committed profile of this workload and plan, used
    # only while the profile's plan matches the one in effect
    traffic = traffic_src = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    wl = f"pagerank_pull_{args.storage}_{args.graph}{args.scale}"
    if os.path.exists(tpath):
        t = json.load(open(tpath)).get(wl)
        if t and t.get("hot_segments") == info.hot_segments and \
                t.get("src_block_bits") == info.src_block_bits and \
                t.get("top_sources") == info.top_sources and \
                t.get("virtual_rows") == info.virtual_rows and \
                t.get("sorted_tiles", 0) == info.sorted_tiles:
            traffic = t["traffic_bytes_per_launch"]
            traffic_src = (f"profiled, not measured in this run: {t['profile']} "
                           f"(rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by a copy kernel)")
    out = {
        "metric": METRIC, "value": round(gteps, 3), "unit": "GTEPS", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.storage == "f32" else f"{args.storage}-storage/f32-accumulate",
        "data": "synthetic",
        "config": {"workload": wl,
                   "graph": gdesc, "num_vertices": n, "num_edges": m,
                   "step": "one PageRank-pull iteration (gather + damping update + L1/active "
                           "reduction + next contribution)"},
        "hbm_gb_per_s": round(nbytes / (ms_per_step * 1e-3) / 1e9, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes": nbytes,
                     "kernel_ms": round(kern_ms, 4),
                     "kernel": ("one iteration = " +
                                ("pr_f32_hot_kernel (hot source segments gathered from LDS) + "
                                 if info.hot_segments else "") +
                                "pr_f32_blk_kernel (source-blocked gather) + "
                                "pr_f32_cmb_kernel (combine + update) + pr_f32_finish"
                                if info.src_block_bits else
                                "one iteration = pr_f32_kernel + pr_f32_finish"),
                     "src_block_bits": info.src_block_bits,
                     "hot_segments": info.hot_segments,
                     "hot_edge_fraction": round(info.hot_edges / max(m, 1), 4),
                     "lds_top_sources": info.top_sources,
                     "lds_top_edge_fraction": round(info.top_edges / max(m, 1), 4),
                     "sorted_tiles": bool(info.sorted_tiles),
                     "measured_copy_gb_per_s": round(copy_gbs, 1)},
    }

    if fixed32 is not None:
        out["fixed32_iteration"] = fixed32
        out["quantized_iteration"] = quant
        out["directions"] = directions
    if fixed32 is not None and not args.no_cpu:
        t0 = time.perf_counter()
        cpu_rows(g, quant, directions)
        log(f"CPU legs of the quantised / push / data-driven rows ({time.perf_counter() - t0:.1f} s)")
    if construction:
        out["graph_construction"] = construction
    if not args.no_ttl:
        t0 = time.perf_counter()
        f64 = None
        if args.f64_iters:
            import oracle
            f64 = oracle.pagerank_f64(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                                      tol=1e-13, max_iterations=args.f64_iters,
                                      threads=HOST_THREADS)
        log(f"FP64 oracle ({time.perf_counter() - t0:.1f} s)")
        out["time_to_l1_1e-6"] = time_to_l1(ag, dg, g, f64)
        log(f"time to L1<=1e-6: {out['time_to_l1_1e-6']} ({time.perf_counter() - t0:.1f} s)")
        dg.close()
        if not args.no_warm:
            t0 = time.perf_counter()
            out["time_to_l1_1e-6"]["learned_warm_start"] = warm_section(ag, dev, g, f64,
                                                                        not args.no_cpu)
            log(f"warm start: {out['time_to_l1_1e-6']['learned_warm_start']} "
                f"({time.perf_counter() - t0:.1f} s)")
    else:
        dg.close()
    if not args.no_batch and args.batch_count > 0:
        t0 = time.perf_counter()
        out["batched"] = batched_section(ag, dev, args.batch_count, 10, 16, not args.no_cpu)
        log(f"batched: {out['batched']} ({time.perf_counter() - t0:.1f} s)")
    if not args.no_ppr:
        t0 = time.perf_counter()
        out["ppr_64"] = ppr_section(ag, dev, args.ppr_scale, 64, not args.no_cpu)
        log(f"ppr: {out['ppr_64']} ({time.perf_counter() - t0:.1f} s)")
    if not args.no_spmv:
        t0 = time.perf_counter()
        out["spmv"] = spmv_section(ag, dev, args.spmv_scale, 20, not args.no_cpu, pkw)
        log(f"spmv: {out['spmv']} ({time.perf_counter() - t0:.1f} s)")
    if not args.no_spmv:
        t0 = time.perf_counter()
        out["config1_rmat16"] = config1_section(ag, dev, not args.no_cpu)
        log(f"config 1: {out['config1_rmat16']} ({time.perf_counter() - t0:.1f} s)")
    if not args.no_er and args.graph == "rmat":
        t0 = time.perf_counter()
        out["er_pagerank"] = er_section(ag, dev, args.scale, 20, not args.no_cpu, pkw, m)
        log(f"er: {out['er_pagerank']} ({time.perf_counter() - t0:.1f} s)")
    dev.close()
    if not args.no_cpu:
        t0 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(g, args.cpu_seconds)
        log(f"cpu baseline {out['cpu_baseline']['value']} GTEPS on "
            f"{out['cpu_baseline']['cores']} cores ({time.perf_counter() - t0:.1f} s)")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
