"""Per-row FP32 bars shared by the parity tests.

The FP32 gather may add a row's terms in any fixed order (tree reductions, source-block
partials), while the reference adds them in sequence in inverse-CSR order
(02_capi_graph/src/algorithms/capi/pageRank.c:2903-2907; the float RTL golden,
03_capi_integration/accelerator_verification/rtl/unit/algorithms/tb/pagerank_kernel_tb.sv:
303-314). Two bars follow from that, both per row and in scaled (unnormalised) units, so the
heaviest rows are held to the same relative standard as the lightest:

* one_step_check -- one iteration from the same start. Any order of adding d terms x_j gives
  a result within gamma_(d-1) * sum|x_j| of the exact sum (gamma_k = k u / (1 - k u),
  u = 2^-24, Higham, Accuracy and Stability of Numerical Algorithms, 2nd ed., eq. 4.4). The
  update base + damp * sum then rounds twice more. So
      |new - (base + damp * exact)| <= damp * gamma_(d-1) * S + 2 ulp(new)  + tiny,
  with the exact sum taken in FP64 over the same FP32 contributions. A dropped term, a term
  added twice or a term from the wrong source moves the row by a whole term and fails
  unless that term is below the bound.
* anchor_check -- after the same number of iterations as an FP64 run from the same start,
  every row is at most twice as far from FP64 as the oracle's sequential FP32 row is, plus
  `ulps` units in the last place of the FP64 value (the order of additions differs, so the
  rounding noise differs; calibrated on the CPU against the oracle run with every row's
  terms reversed, which needs up to 40 ulp at RMAT-20).
"""
from __future__ import annotations

import numpy as np

U = 2.0 ** -24
DAMP = np.float32(0.85)
BASE = np.float32(1.0) - DAMP


def contributions(out_degree, scaled) -> np.ndarray:
    """rank / out_degree in FP32 (u32 -> float conversion first), 0 for sinks
    (pageRank.c:1026-1033)."""
    od = np.asarray(out_degree)
    r = np.asarray(scaled, np.float32)
    c = np.zeros(r.size, np.float32)
    nz = od > 0
    c[nz] = r[nz] / od[nz].astype(np.float32)
    return c


def exact_row_sums(inv_idx, in_degree, inv_src, vals):
    """Row sums of vals[src] in FP64 (exact up to 2^-53 relative) and sums of |vals[src]|."""
    deg = np.asarray(in_degree, np.int64)
    n = deg.size
    src = np.asarray(inv_src)
    m = int(deg.sum())
    if m == 0:
        return np.zeros(n), np.zeros(n)
    # edges of row v are inv_src[inv_idx[v] : inv_idx[v] + deg[v]]
    idx = np.asarray(inv_idx, np.int64)
    if np.array_equal(idx[1:], (idx + deg)[:-1]) and idx[0] == 0:
        gathered = np.asarray(vals, np.float64)[src[:m]]
    else:
        rows = np.repeat(np.arange(n), deg)
        pos = np.arange(m) - np.repeat(np.cumsum(deg) - deg, deg) + idx[rows]
        gathered = np.asarray(vals, np.float64)[src[pos]]
    row = np.repeat(np.arange(n), deg)
    s = np.bincount(row, weights=gathered, minlength=n)
    a = np.bincount(row, weights=np.abs(gathered), minlength=n)
    return s, a


def one_step_bound(g_out_degree, inv_idx, in_degree, inv_src, init_scaled):
    """(expected, bound): the exact one-iteration value of every row and the largest distance
    an FP32 evaluation that adds the row's terms in any order may have from it."""
    c = contributions(g_out_degree, init_scaled)
    s, a = exact_row_sums(inv_idx, in_degree, inv_src, c)
    d = np.asarray(in_degree, np.float64)
    k = np.maximum(d - 1.0, 0.0)
    gamma = k * U / (1.0 - k * U)
    expected = float(BASE) + float(DAMP) * s
    ulp = np.spacing(np.abs(expected).astype(np.float32)).astype(np.float64)
    bound = float(DAMP) * gamma * a * (1.0 + 4 * U) + 2.0 * ulp + 1e-300
    return expected, bound


def one_step_check(g, init_scaled, got_scaled, what="gpu"):
    """Asserts the one-iteration bar on every row; returns the largest bound fraction used."""
    expected, bound = one_step_bound(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                                     init_scaled)
    diff = np.abs(np.asarray(got_scaled, np.float64) - expected)
    bad = np.nonzero(diff > bound)[0]
    assert bad.size == 0, (
        f"{what}: {bad.size} rows outside the reordered-sum bound; first rows "
        f"{bad[:5].tolist()} in-degree {np.asarray(g.in_degree)[bad[:5]].tolist()} "
        f"diff {diff[bad[:5]].tolist()} bound {bound[bad[:5]].tolist()}")
    return float((diff / bound).max()) if diff.size else 0.0


def anchor_violations(got, ora, f64, ulps=64):
    got = np.asarray(got, np.float64)
    ora = np.asarray(ora, np.float64)
    f64 = np.asarray(f64, np.float64)
    ulp = np.spacing(np.abs(f64).astype(np.float32)).astype(np.float64)
    lim = 2.0 * np.abs(ora - f64) + ulps * ulp + 1e-300
    return np.nonzero(np.abs(got - f64) > lim)[0]


def anchor_check(got, ora, f64, ulps=64, what="gpu"):
    bad = anchor_violations(got, ora, f64, ulps)
    assert bad.size == 0, (
        f"{what}: {bad.size} rows farther from FP64 than twice the oracle plus {ulps} ulp; "
        f"first rows {bad[:5].tolist()} got {np.asarray(got)[bad[:5]].tolist()} "
        f"oracle {np.asarray(ora)[bad[:5]].tolist()} f64 {np.asarray(f64)[bad[:5]].tolist()}")


def f32_row_bars(dg, g, iterations, threads=8, seed=0):
    """Both bars on a graph uploaded to the GPU (dg, an accelgraph_amd.DeviceGraph): one
    iteration from a random positive start against the reordered-sum bound, then `iterations`
    iterations from the uniform start against FP64 run for as many iterations. Returns the
    GPU run of the second check."""
    import accelgraph_amd as ag
    import oracle
    init = (np.random.default_rng(seed).random(g.n) * 2 + 0.01).astype(np.float32)
    one = dg.pagerank(max_iterations=1, stop_rule=ag.STOP_NONE, init_ranks=init, scaled=True,
                      history=False)
    used = one_step_check(g, init, one.scaled)
    run = dg.pagerank(max_iterations=iterations, stop_rule=ag.STOP_NONE, init=ag.INIT_UNIFORM,
                      history=False)
    ora = oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                               max_iterations=iterations, stop_rule=2, init=1, threads=threads)
    f64 = oracle.pagerank_f64(g.out_degree, g.inv_idx, g.in_degree, g.inv_src, tol=0.0,
                              max_iterations=iterations, threads=threads)
    anchor_check(run.ranks, ora.ranks, f64)
    rel = lambda r: float(np.max(np.abs(r - f64) / np.maximum(f64, 1e-300)))
    print(f"  row bars: one step uses {used:.3f} of the bound; after {iterations} iterations "
          f"max |gpu-f64|/f64 {rel(run.ranks):.2e} (oracle {rel(ora.ranks):.2e})", flush=True)
    return run
