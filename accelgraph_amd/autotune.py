"""Autotuner: picks the pull plan's bin thresholds and launch shape for a graph, the part
size and heavy-row threshold of a many-graph batch, and the occupancy of the 64-column PPR.

Coordinate descent over a small grid (source blocking, bin thresholds, occupancy, row order):
starting from the defaults, each knob in turn is set to every value of its grid while the
others stay fixed, each candidate plan is timed on the GPU by
ag_autotune_plan (median of `repeats` runs of `iterations` FP32 iterations), and the fastest
value is kept; two passes. Results are cached in a JSON table keyed by a graph signature
(vertex and edge counts plus the log2 histogram of in-degrees), so the same graph shape is
tuned once. A plan changes only the schedule: integer results are identical under every
candidate and FP32 results differ only by summation order.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os

import numpy as np

from . import _abi, _check, lib, plan_params

# a plan holds at most this many source blocks (AG_MAX_SRC_BLOCKS of the planner)
MAX_SRC_BLOCKS = 256

GRID = {
    "src_block_bits": [255, 18, 19, 20],   # 255: no source blocking
    "blocked_min_bin": [0, 1],
    "thread_max_deg": [8, 16, 32],
    "group16_max_deg": [64, 128, 256],
    "wave_max_deg": [1024, 2048, 4096],
    "block_max_deg": [8192, 16384, 65536],
    "wg_per_cu": [3, 4, 5, 6, 8],
    "order": [0, 1],
    "hot_segments": [0, 32, 128],
    "top_table": [_abi.TOP_AUTO, _abi.TOP_OFF, _abi.TOP_ON],
    "tile_order": [_abi.TILES_AUTO, _abi.TILES_ROUND_ROBIN, _abi.TILES_TICKET],
}

# the batched forms: ag_batch_options fields and ag_ppr_options.wg_per_cu
BATCH_GRID = {
    "part_edges": [1024, 2048, 4096, 8192, 16384],
    "heavy_deg": [16, 32, 64],
}
PPR_GRID = {
    "wg_per_cu": [2, 4, 6, 8],
}


def default_table_path() -> str:
    env = os.environ.get("ACCELGRAPH_TUNING_TABLE")
    if env:
        return env
    base = os.environ.get("XDG_CACHE_HOME", os.path.join(os.path.expanduser("~"), ".cache"))
    return os.path.join(base, "accelgraph_amd", "tuning.json")


def signature(n: int, m: int, in_degree: np.ndarray) -> str:
    deg = np.asarray(in_degree, dtype=np.uint64)
    buckets = np.bincount(np.floor(np.log2(deg + 1)).astype(np.int64), minlength=33)
    h = hashlib.sha1(buckets.astype(np.int64).tobytes()).hexdigest()[:16]
    return f"n{n}-m{m}-{h}"


def _load(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path: str, table: dict) -> None:
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _time(dev, arrays, cands, iterations, repeats):
    n, od, ii, dg, src = arrays
    arr = (_abi.PlanParams * len(cands))(*cands)
    ms = np.zeros(len(cands), np.float64)
    best = C.c_uint32(0)
    u32p = C.POINTER(C.c_uint32)
    _check(lib().ag_autotune_plan(dev.handle, n, src.size, od.ctypes.data_as(u32p),
                                  ii.ctypes.data_as(u32p), dg.ctypes.data_as(u32p),
                                  src.ctypes.data_as(u32p), arr, len(cands), iterations, repeats,
                                  ms.ctypes.data_as(C.POINTER(C.c_double)), C.byref(best)),
           "ag_autotune_plan")
    return ms, best.value


def _params_dict(p: _abi.PlanParams) -> dict:
    return {name: getattr(p, name) for name, _ in _abi.PlanParams._fields_}


def autotune(dev, graph, iterations: int = 5, repeats: int = 3, passes: int = 2,
             table_path: str | None = None, force: bool = False):
    """Returns (PlanParams, report). report["cached"] tells whether the table answered."""
    path = table_path or default_table_path()
    sig = signature(graph.n, graph.m, graph.in_degree)
    table = _load(path)
    if not force and sig in table:
        return plan_params(**table[sig]["params"]), dict(table[sig], cached=True)
    arrays = (graph.n, np.ascontiguousarray(graph.out_degree, np.uint32),
              np.ascontiguousarray(graph.inv_idx, np.uint32),
              np.ascontiguousarray(graph.in_degree, np.uint32),
              np.ascontiguousarray(graph.inv_src, np.uint32))
    cur = _params_dict(plan_params())
    base_ms, _ = _time(dev, arrays, [plan_params(**cur)], iterations, repeats)
    best_ms = float(base_ms[0])
    trials = [(dict(cur), best_ms)]
    for _ in range(passes):
        for knob, values in GRID.items():
            cands, dicts = [], []
            for v in values:
                d = dict(cur, **{knob: v})
                if not (d["thread_max_deg"] <= d["group16_max_deg"] <= d["wave_max_deg"]
                        <= d["block_max_deg"]):
                    continue
                if (knob == "src_block_bits" and v != 255
                        and (graph.n + (1 << v) - 1) >> v > MAX_SRC_BLOCKS):
                    continue   # more source blocks than a plan allows
                dicts.append(d)
                cands.append(plan_params(**d))
            ms, b = _time(dev, arrays, cands, iterations, repeats)
            trials += [(d, float(t)) for d, t in zip(dicts, ms)]
            if float(ms[b]) < best_ms:
                best_ms, cur = float(ms[b]), dicts[b]
    entry = {"params": cur, "ms_per_iteration": best_ms,
             "default_ms_per_iteration": float(base_ms[0]), "trials": len(trials)}
    table[sig] = entry
    _save(path, table)
    return plan_params(**cur), dict(entry, cached=False)


def _descend(grid: dict, start: dict, timer, passes: int):
    """Coordinate descent: every knob in turn over its grid values (others fixed), keeping
    the fastest. timer(dict) -> ms. Returns (best dict, best ms, default ms, trials)."""
    cur = dict(start)
    base = timer(cur)
    best = base
    trials = 1
    for _ in range(passes):
        for knob, values in grid.items():
            for v in values:
                if cur.get(knob) == v:
                    continue
                cand = dict(cur, **{knob: v})
                t = timer(cand)
                trials += 1
                if t < best:
                    best, cur = t, cand
    return cur, best, base, trials


def batch_signature(batch) -> str:
    sizes = np.diff(batch.edge_off.astype(np.int64))
    h = hashlib.sha1(np.sort(sizes).tobytes()).hexdigest()[:16]
    return f"batch-G{batch.vert_off.size - 1}-n{int(batch.vert_off[-1])}-m{int(batch.edge_off[-1])}-{h}"


def autotune_batch(batch, repeats: int = 3, passes: int = 1, table_path: str | None = None,
                   force: bool = False, **solve):
    """Part size and heavy-row threshold of a Batch, timed on the solve itself (median device
    time of `repeats` captured solves with the caller's options, default L1 <= 1e-6 from the
    uniform start). Returns (options dict, report)."""
    path = table_path or default_table_path()
    sig = batch_signature(batch)
    table = _load(path)
    if not force and sig in table:
        return dict(table[sig]["options"]), dict(table[sig], cached=True)
    kw = dict(dict(max_iterations=300, stop_rule=_abi.STOP_L1, l1_tolerance=1e-6,
                   init=_abi.INIT_UNIFORM), **solve)

    def timer(opts):
        ms = []
        for _ in range(repeats + 1):             # the first run captures the solve
            ms.append(batch.pagerank(use_hip_graph=1, **dict(kw, **opts))[3])
        return float(np.median(ms[1:]))

    start = {"part_edges": 4096, "heavy_deg": 32}
    best, best_ms, base_ms, trials = _descend(BATCH_GRID, start, timer, passes)
    entry = {"options": best, "ms": best_ms, "default_ms": base_ms, "default": start,
             "trials": trials}
    table[sig] = entry
    _save(path, table)
    return dict(best), dict(entry, cached=False)


def autotune_ppr(dg, seeds, iterations: int = 8, repeats: int = 3, passes: int = 1,
                 storage: int = _abi.STORE_BF16, table_path: str | None = None,
                 force: bool = False):
    """Occupancy of the PPR iteration kernel for one graph and column count: each candidate
    times `iterations` fixed iterations of a captured solve (median of `repeats`).
    Returns (options dict, report)."""
    path = table_path or default_table_path()
    info = dg.info()
    sig = f"ppr-k{len(seeds)}-s{storage}-n{info.num_vertices}-m{info.num_edges}"
    table = _load(path)
    if not force and sig in table:
        return dict(table[sig]["options"]), dict(table[sig], cached=True)

    def timer(opts):
        ms = []
        for _ in range(repeats + 1):
            _, _, r = dg.ppr(seeds, max_iterations=iterations, stop_rule=_abi.STOP_NONE,
                             storage=storage, use_hip_graph=1, **opts)
            ms.append(r.loop_ms / max(r.executed, 1))
        return float(np.median(ms[1:]))

    start = {"wg_per_cu": 0}
    best, best_ms, base_ms, trials = _descend(PPR_GRID, start, timer, passes)
    entry = {"options": best, "ms_per_iteration": best_ms, "default_ms_per_iteration": base_ms,
             "default": start, "trials": trials}
    table[sig] = entry
    _save(path, table)
    return dict(best), dict(entry, cached=False)
