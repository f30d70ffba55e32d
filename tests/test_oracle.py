"""Pins the CPU oracle (oracle/ag_oracle.c) before anything is compared against it.

Checks, each chosen so that a plausible slip (dropped term, wrong sign/index, transposed
operand) fails one of them:
  * the reference's RTL kernel known-answer fixtures (tests/golden/rtl_gather_kat.json, made
    with exact arithmetic by tests/golden/make_rtl_kat.py) for the gather arithmetic;
  * closed forms of PageRank on the chain / star / cycle / isolated-vertex graphs;
  * brute-force dense FP64 power iteration (numpy matrices) on tiny random graphs;
  * exact rational definitions of the fixed-point encoders.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
from golden.make_rtl_kat import f32_round

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "rtl_gather_kat.json")))


def _csr(n, edges):
    """inverse CSR in edge-list order (the RTL fixtures' own convention)"""
    rows = [[] for _ in range(n)]
    for s, d in edges:
        rows[d].append(s)
    idx, deg, src = [], [], []
    for v in range(n):
        idx.append(len(src))
        deg.append(len(rows[v]))
        src.extend(rows[v])
    outdeg = np.zeros(n, np.uint32)
    for s, _ in edges:
        outdeg[s] += 1
    return outdeg, np.array(idx, np.uint32), np.array(deg, np.uint32), np.array(src, np.uint32)


# ------------------------------------------------------------------ RTL known answers

@pytest.mark.parametrize("case", [c for c in KAT["cases"] if c["op"] == "pagerank_gather"],
                         ids=lambda c: f"{c['kind']}-{c['fixture']}-{c['cls']}")
def test_oracle_pagerank_gather_kat(case):
    n = case["n"]
    if n == 0:
        return
    m = len(case["values"])
    idx = np.array(case["idx"], np.uint32)
    deg = np.array(case["deg"], np.uint32)
    src = np.arange(m, dtype=np.uint32)  # edge j reads its own value
    if case["kind"] == "float":
        vals = np.array([int(v) for v in case["values"]], np.uint32).view(np.float32)
        got = oracle.gather_f32(idx, deg, src, vals).view(np.uint32)
        exp = np.array([int(v) for v in case["expected"]], np.uint32)
    else:
        vals = np.array([int(v) for v in case["values"]], np.uint64)
        got = oracle.gather_u64(idx, deg, src, vals)
        exp = np.array([int(v) for v in case["expected"]], np.uint64)
    np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("case", [c for c in KAT["cases"] if c["op"] == "spmv_gather"],
                         ids=lambda c: f"{c['kind']}-{c['fixture']}-{c['cls']}")
def test_oracle_spmv_kat(case):
    n = case["n"]
    if n == 0:
        return
    m = len(case["x"])
    idx = np.array(case["idx"], np.uint32)
    deg = np.array(case["deg"], np.uint32)
    src = np.arange(m, dtype=np.uint32)
    if case["kind"] == "float":
        x = np.array([int(v) for v in case["x"]], np.uint32).view(np.float32)
        w = np.array([int(v) for v in case["w"]], np.uint32).view(np.float32)
        got = oracle.spmv_f32(idx, deg, src, w, x).view(np.uint32)
        exp = np.array([int(v) for v in case["expected"]], np.uint32)
    else:
        x = np.array([int(v) for v in case["x"]], np.uint32)
        w = np.array([int(v) for v in case["w"]], np.uint32)
        got = oracle.spmv_fixed(idx, deg, src, w, 0, x)
        exp = np.array([int(v) for v in case["expected"]], np.uint64)
    np.testing.assert_array_equal(got, exp)


def test_kat_covers_reference_run_lists():
    pr = {(c["kind"], c["fixture"], c["cls"]) for c in KAT["cases"] if c["op"] == "pagerank_gather"}
    # pagerank_kernel_tb.sv:617-634 drives these
    for need in [("float", "k4", "max"), ("fixed64", "k4", "mixed"), ("quant32", "k4", "max"),
                 ("float", "directed-chain", "mixed"), ("fixed64", "dangling-sink", "zero")]:
        assert need in pr
    # the FP32 overflow case saturates to +inf exactly as binary32 adds do
    k4max = [c for c in KAT["cases"] if c["op"] == "pagerank_gather" and c["kind"] == "float"
             and c["fixture"] == "k4" and c["cls"] == "max"][0]
    assert all(int(e) == 0x7F800000 for e in k4max["expected"])


# ------------------------------------------------------------------ closed forms

def _run(n, edges, **kw):
    od, idx, deg, src = _csr(n, edges)
    return oracle.pagerank_pull(od, idx, deg, src, **kw)


def test_chain_closed_form():
    r = _run(4, [(0, 1), (1, 2), (2, 3)], max_iterations=100)
    b = np.float32(1.0) - np.float32(0.85)
    r1 = b + np.float32(0.85) * b
    r2 = b + np.float32(0.85) * r1
    r3 = b + np.float32(0.85) * r2
    assert abs(float(r1) - 0.2775) < 1e-6 and abs(float(r3) - 0.47799375) < 1e-6
    np.testing.assert_allclose(r.scaled, [b, r1, r2, r3], rtol=0, atol=0)
    assert r.iterations == 3 and r.executed == 4 and r.active == 0


def test_star_closed_form():
    r = _run(5, [(0, 1), (0, 2), (0, 3), (0, 4)], max_iterations=100)
    assert r.iterations == 1
    np.testing.assert_allclose(r.scaled[1:], 0.15 + 0.85 * 0.15 / 4, rtol=1e-6)
    np.testing.assert_allclose(r.scaled[0], 0.15, rtol=1e-6)
    np.testing.assert_allclose(r.ranks, r.scaled / 5, rtol=1e-7)


def test_cycle_closed_form():
    # from r0 = 0.15, iteration it gives |delta| = 0.15 * 0.85^(it+1); with eps = 1e-3 the
    # first quiet iteration is it+1 = 31, and r = 1 - 0.85^(it+2)
    r = _run(3, [(0, 1), (1, 2), (2, 0)], max_iterations=200, epsilon=1e-3)
    it = r.iterations
    assert 0.15 * 0.85 ** (it + 1) < 1e-3 <= 0.15 * 0.85 ** it
    assert it == 30
    np.testing.assert_allclose(r.scaled, 1 - 0.85 ** (it + 2), rtol=1e-6)
    # converged to the fixed point 1 (scaled), i.e. 1/N normalised
    r = _run(3, [(0, 1), (1, 2), (2, 0)], max_iterations=400, epsilon=1e-9)
    np.testing.assert_allclose(r.ranks, 1 / 3, rtol=1e-6)


def test_isolated_and_sources_get_base():
    r = _run(4, [(0, 1)], max_iterations=1, stop_rule=2)
    np.testing.assert_allclose(r.scaled[[0, 2, 3]], np.float32(1) - np.float32(0.85), rtol=0)


def test_dangling_mass_is_dropped():
    # 0 -> 1, 1 is a sink: the rank sum stays below 1 (pageRank.c:1029-1032)
    r = _run(2, [(0, 1)], max_iterations=100)
    assert r.rank_sum < 1.0
    np.testing.assert_allclose(r.scaled, [0.15, 0.15 + 0.85 * 0.15], rtol=1e-6)


def _dense_f64(n, edges, iters, init):
    A = np.zeros((n, n))
    od = np.zeros(n)
    for s, d in edges:
        A[d, s] += 1.0
        od[s] += 1
    P = A * np.where(od > 0, 1.0 / np.maximum(od, 1), 0.0)[None, :]
    r = np.full(n, init)
    for _ in range(iters):
        r = 0.15 + 0.85 * P @ r
    return r


@pytest.mark.parametrize("seed", range(6))
def test_bruteforce_dense_power_iteration(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 40))
    m = int(rng.integers(0, 4 * n))
    edges = sorted({(int(rng.integers(n)), int(rng.integers(n))) for _ in range(m)})
    edges = [(s, d) for s, d in edges if s != d]
    r = _run(n, edges, max_iterations=25, stop_rule=2)
    dense = _dense_f64(n, edges, 25, 0.15)
    np.testing.assert_allclose(r.scaled, dense, rtol=2e-6)
    od, idx, deg, src = _csr(n, edges)
    f64 = oracle.pagerank_f64(od, idx, deg, src, tol=1e-15, max_iterations=5000)
    np.testing.assert_allclose(f64, _dense_f64(n, edges, 400, 1.0) / n, rtol=1e-9)


@pytest.mark.parametrize("seed", range(6))
def test_ppr_f64_matches_linear_solve(seed):
    """Personalised PageRank fixed point in closed form: r = (I - 0.85 P)^-1 p, with p the
    restart vector (0.15 n on the seed vertex, or 0.15 everywhere for the uniform restart) and
    P the column-stochastic transition matrix with dangling columns zero (pageRank.c:1026-1033).
    The oracle iterates; the solve does not."""
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(2, 48))
    m = int(rng.integers(0, 4 * n))
    edges = sorted({(int(rng.integers(n)), int(rng.integers(n))) for _ in range(m)})
    edges = [(s, d) for s, d in edges if s != d]
    A = np.zeros((n, n))
    od = np.zeros(n)
    for s, d in edges:
        A[d, s] += 1.0
        od[s] += 1
    P = A * np.where(od > 0, 1.0 / np.maximum(od, 1), 0.0)[None, :]
    M = np.eye(n) - 0.85 * P
    odeg, idx, deg, src = _csr(n, edges)
    for sv in [0xFFFFFFFF, int(rng.integers(n)), n - 1]:
        p = np.full(n, 0.15) if sv == 0xFFFFFFFF else 0.15 * n * (np.arange(n) == sv)
        want = np.linalg.solve(M, p) / n
        got = oracle.ppr_f64(odeg, idx, deg, src, sv, tol=1e-15, max_iterations=20000)
        # atol: what the iteration leaves of its all-ones start where the exact answer is 0
        np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12)
    # the uniform restart is plain PageRank
    np.testing.assert_allclose(oracle.ppr_f64(odeg, idx, deg, src, 0xFFFFFFFF, tol=1e-15,
                                              max_iterations=20000),
                               oracle.pagerank_f64(odeg, idx, deg, src, tol=1e-15,
                                                   max_iterations=20000), rtol=1e-12)


@pytest.mark.parametrize("seed", range(4))
def test_ppr_residual_against_linear_solve(seed):
    """The residual p + 0.85 P r - r (normalised units) of the exact solution (linear solve)
    is at the rounding floor; of r = 0 it is p exactly (L1 = 0.15 for every restart); of a
    perturbed solution r* + delta it is (I - 0.85 P) delta, so it lies within
    [0.15, 1.85] * ||delta||_1 and bounds ||r - r*||_1 <= residual / 0.15. Several columns at
    once, vertex-major."""
    rng = np.random.default_rng(300 + seed)
    n = int(rng.integers(3, 60))
    m = int(rng.integers(n, 5 * n))
    edges = sorted({(int(rng.integers(n)), int(rng.integers(n))) for _ in range(m)})
    edges = [(s, d) for s, d in edges if s != d]
    A = np.zeros((n, n))
    od = np.zeros(n)
    for s, d in edges:
        A[d, s] += 1.0
        od[s] += 1
    P = A * np.where(od > 0, 1.0 / np.maximum(od, 1), 0.0)[None, :]
    M = np.eye(n) - 0.85 * P
    odeg, idx, deg, src = _csr(n, edges)
    seeds = np.array([0xFFFFFFFF, 0, n - 1, int(rng.integers(n))], np.uint32)
    cols = []
    for sv in seeds:
        p = np.full(n, 0.15 / n) if sv == 0xFFFFFFFF else 0.15 * (np.arange(n) == sv)
        cols.append(np.linalg.solve(M, p))
    exact = np.stack(cols, 1)
    res, nrm = oracle.ppr_residual(odeg, idx, deg, src, seeds, exact.astype(np.float32))
    # FP32 storage of the exact solution leaves a residual of a few FP32 ulps of the column
    assert np.all(res <= 8 * 2.0 ** -24 * nrm)
    np.testing.assert_allclose(nrm, np.abs(exact).sum(0), rtol=1e-6)
    res0, nrm0 = oracle.ppr_residual(odeg, idx, deg, src, seeds, np.zeros((n, 4), np.float32))
    np.testing.assert_allclose(res0, 0.15, rtol=1e-12)
    assert np.all(nrm0 == 0)
    delta = (rng.normal(size=(n, 4)) * 1e-3 / n).astype(np.float32)
    r = (exact + delta).astype(np.float32)
    res, _ = oracle.ppr_residual(odeg, idx, deg, src, seeds, r)
    d1 = np.abs(r.astype(np.float64) - exact).sum(0)
    want = np.abs(M @ (r.astype(np.float64) - exact)).sum(0)
    np.testing.assert_allclose(res, want, rtol=1e-6, atol=1e-9 / n)
    assert np.all(d1 <= res / 0.15 * (1 + 1e-6) + 1e-9 / n)


def test_fixed_encode_array_equals_scalar():
    rng = np.random.default_rng(8)
    x = np.concatenate([np.array([0.0, -1.0, np.nan, np.inf, 0.5, 65535.996, 1e30], np.float32),
                        (rng.random(5000) * 2.0 ** rng.integers(-20, 20, 5000)).astype(np.float32)])
    for bits in (64, 32, 16, 8):
        np.testing.assert_array_equal(oracle.fixed_encode_array(bits, x),
                                      oracle.fixed_encode(bits, x))


def test_openmp_matches_sequential():
    rng = np.random.default_rng(3)
    n = 3000
    edges = {(int(rng.integers(n)), int(rng.integers(n))) for _ in range(20000)}
    edges = sorted((s, d) for s, d in edges if s != d)
    od, idx, deg, src = _csr(n, edges)
    a = oracle.pagerank_pull(od, idx, deg, src, max_iterations=50)
    b = oracle.pagerank_pull(od, idx, deg, src, max_iterations=50, threads=4)
    np.testing.assert_array_equal(a.ranks, b.ranks)
    assert a.iterations == b.iterations
    assert abs(a.error_total - b.error_total) <= 1e-12


# ------------------------------------------------------------------ fixed point

def _exact_code(x, frac, width):
    if not (x > 0) or math.isnan(x):
        return 0
    if math.isinf(x):
        return (1 << width) - 1
    q = math.floor(Fraction(float(x)) * (1 << frac) + Fraction(1, 2))
    return min(q, (1 << width) - 1)


FIX_CASES = [0.0, -1.0, float("nan"), float("inf"), 1.0, 0.5, 2.0 ** -17, 2.0 ** -17 * 0.999,
             2.0 ** -18, 0.15, 0.85, 1 / 3, 1e-7, 65535.99609375, 65536.0, 1e30, 1e-30, 255.9,
             256.0, 15.97, 16.0, 3.0517578125e-05, 2.0 ** 31, 2.0 ** 32, 1.5 * 2 ** 31]


@pytest.mark.parametrize("bits,frac", [(64, 32), (32, 16), (16, 8), (8, 4)])
def test_fixed_encode_exact(bits, frac):
    xs = np.array(FIX_CASES, np.float32)
    rng = np.random.default_rng(bits)
    xs = np.concatenate([xs, rng.random(500).astype(np.float32) * 4,
                         (rng.random(200) * 2.0 ** rng.integers(-30, 40, 200)).astype(np.float32)])
    got = oracle.fixed_encode(bits, xs)
    exp = np.array([_exact_code(float(x), frac, bits) for x in xs], np.uint64)
    np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("bits,frac", [(64, 32), (32, 16), (16, 8), (8, 4)])
def test_fixed_decode_exact(bits, frac):
    qs = np.array([0, 1, 2, 3, 65536, 65537, 2 ** 24 + 1, 2 ** 40 + 12345, 2 ** 63 + 1,
                   2 ** 64 - 1], np.uint64)
    got = oracle.fixed_decode(bits, qs)
    for q, g in zip(qs, got):
        exact = Fraction(int(q), 1 << frac)
        if bits == 64:
            assert g == float(exact)  # one round-to-nearest-even to double
        else:
            # round the integer to binary32 (nearest even), then scale exactly
            f32 = np.uint32(f32_round(Fraction(int(q)))).view(np.float32)
            assert g == float(f32) / (1 << frac)


# ------------------------------------------------------------------ data-driven pull

def _dd(g, **kw):
    return oracle.pagerank_dd_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                                   g.edges_idx, g.out_degree, g.fwd_dest, **kw)


def test_dd_pull_with_zero_epsilon_is_the_pull():
    """With epsilon 0 every listed vertex takes its new rank and lists its out-neighbours. A
    vertex left off the list has in-neighbours that did not change since it was last
    evaluated, so re-evaluating it would give its current rank: the ranks equal the plain
    pull's bit for bit after every number of iterations."""
    import accelgraph_amd as ag
    g = ag.Graph.rmat(11)
    for k in (1, 2, 5, 17):
        d = _dd(g, max_iterations=k, epsilon=0.0)
        p = oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                                 max_iterations=k, stop_rule=2)
        np.testing.assert_array_equal(d.scaled.view(np.uint32), p.scaled.view(np.uint32))


def test_dd_pull_star_and_cycle():
    import accelgraph_amd as ag
    # star: the centre (no in-edges) never changes, the leaves change once, then nothing is
    # listed: stop at iteration 1 with the pull's ranks
    g = ag.Graph.from_edges([0, 0, 0, 0], [1, 2, 3, 4])
    d = _dd(g, max_iterations=100)
    assert d.iterations == 1 and d.active == 0
    np.testing.assert_allclose(d.scaled[1:], 0.15 + 0.85 * 0.15 / 4, rtol=1e-6)
    assert d.scaled[0] == np.float32(1) - np.float32(0.85)
    # cycle: every vertex sees the same delta, so all update until the delta drops below eps;
    # the last (sub-epsilon) update is not applied
    g = ag.Graph.from_edges([0, 1, 2], [1, 2, 0])
    d = _dd(g, max_iterations=200, epsilon=1e-3)
    p = oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                             max_iterations=200, epsilon=1e-3)
    assert d.iterations == p.iterations == 30
    np.testing.assert_allclose(d.scaled, 1 - 0.85 ** (d.iterations + 1), rtol=1e-6)


def test_dd_pull_final_ranks_are_epsilon_fixed_points():
    """When the worklist empties, re-evaluating any vertex moves it by less than epsilon: a
    listed vertex was re-evaluated in the last iteration and did not move by epsilon, and an
    unlisted one has unchanged in-neighbours since its last evaluation."""
    import accelgraph_amd as ag
    g = ag.Graph.rmat(12)
    for eps in (1e-3, 1e-5):
        d = _dd(g, max_iterations=500, epsilon=eps)
        assert d.executed == d.iterations + 1 and d.active == 0
        one = oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                                   max_iterations=1, stop_rule=2, init=2,
                                   init_ranks=d.scaled)
        assert np.abs(one.scaled.astype(np.float64) - d.scaled).max() < eps
        f64 = oracle.pagerank_f64(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                                  tol=1e-15, max_iterations=3000)
        # e = r + 0.85 P e with |r_v| < eps and ||P||_1 <= 1 (dangling mass is dropped):
        # sum |e| <= N eps / 0.15 in scaled units, i.e. eps / 0.15 normalised
        assert np.abs(d.ranks.astype(np.float64) - f64).sum() <= eps / 0.15 * 1.001


# ------------------------------------------------------------------ push, parallel forms

def _push(g, **kw):
    return oracle.pagerank_push(g.out_degree, g.edges_idx, g.out_degree, g.fwd_dest, **kw)


@pytest.mark.parametrize("kind", [0, 64, 32, 16, 8])
def test_push_sequential_equals_pull(kind):
    """Scattering sources in ascending order adds each vertex's in-edges in the inverse CSR's
    order (rows sorted by source), so the sequential push is the pull bit for bit."""
    import accelgraph_amd as ag
    g = ag.Graph.rmat(11)
    p = oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src, kind=kind,
                             max_iterations=80, epsilon=1e-4)
    q = _push(g, kind=kind, max_iterations=80, epsilon=1e-4)
    assert (q.iterations, q.executed) == (p.iterations, p.executed)
    np.testing.assert_array_equal(q.scaled.view(np.uint32), p.scaled.view(np.uint32))
    np.testing.assert_array_equal(q.hist_active, p.hist_active)


def test_push_threads():
    """With threads the fixed-point sums stay exact (the pull's ranks bit for bit); FP32 adds
    in the threads' order and stays within rounding."""
    import accelgraph_amd as ag
    g = ag.Graph.rmat(12)
    p = oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src, kind=32,
                             max_iterations=40, stop_rule=2)
    q = _push(g, kind=32, max_iterations=40, stop_rule=2, threads=4)
    np.testing.assert_array_equal(q.scaled.view(np.uint32), p.scaled.view(np.uint32))
    p = oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                             max_iterations=40, stop_rule=2)
    q = _push(g, max_iterations=40, stop_rule=2, threads=4)
    np.testing.assert_allclose(q.scaled, p.scaled, rtol=2e-6)


def test_dd_pull_threads_equal_sequential():
    import accelgraph_amd as ag
    g = ag.Graph.rmat(12)
    a = _dd(g, max_iterations=300, epsilon=1e-5)
    b = _dd(g, max_iterations=300, epsilon=1e-5, threads=4)
    assert (a.iterations, a.executed) == (b.iterations, b.executed)
    np.testing.assert_array_equal(a.scaled.view(np.uint32), b.scaled.view(np.uint32))
    np.testing.assert_array_equal(a.hist_active, b.hist_active)
    np.testing.assert_allclose(a.hist_err, b.hist_err, rtol=1e-12)


# ------------------------------------------------------------------ SpMV push

def test_mul_fixed64_is_the_q32_32_product():
    """MULFixed64V1 as defined here: bits [32, 96) of the exact 128-bit product."""
    rng = np.random.default_rng(4)
    vals = [0, 1, (1 << 32), (1 << 32) + 1, (1 << 64) - 1, 0x123456789ABCDEF0,
            *rng.integers(0, 2 ** 63, 40, dtype=np.uint64).tolist()]
    for a in vals:
        for b in vals[:12]:
            assert oracle.mul_fixed64(a, b) == ((a * b) >> 32) % (1 << 64)


@pytest.mark.parametrize("weighted", [False, True])
def test_spmv_push_sequential(weighted):
    """Sequentially the push adds each y[u]'s terms in ascending source order, the inverse
    CSR's: one repetition equals the pull bit for bit. Repetitions accumulate (fixed point:
    exactly reps times one product; FP32 within rounding), and the fixed and FP32 results
    agree to the Q32.32 resolution."""
    import accelgraph_amd as ag
    g = ag.Graph.rmat(11)
    rng = np.random.default_rng(8)
    if weighted:
        g.set_weights(rng.random(g.m).astype(np.float32))
    fw, iw = g.fwd_weight, g.inv_weight
    x = rng.random(g.n).astype(np.float32)
    fwd = (g.edges_idx, g.out_degree, g.fwd_dest)
    y1 = oracle.spmv_push_f32(*fwd, fw, x)
    pull = oracle.spmv_f32(g.inv_idx, g.in_degree, g.inv_src, iw, x)
    np.testing.assert_array_equal(y1.view(np.uint32), pull.view(np.uint32))
    y3 = oracle.spmv_push_f32(*fwd, fw, x, reps=3)
    # 3d float additions of positive terms: relative error below about 3d * 2^-24
    bound = (3 * g.in_degree.astype(np.float64) + 3) * 2.0 ** -24 * 3 * np.abs(y1) + 1e-30
    assert np.all(np.abs(y3 - 3 * y1.astype(np.float64)) <= bound)
    np.testing.assert_array_equal(oracle.spmv_push_f32(*fwd, fw, x, reps=0), np.zeros(g.n))
    xq = oracle.fixed_encode(64, x)
    q1 = oracle.spmv_push_fixed(*fwd, fw, xq)
    q3 = oracle.spmv_push_fixed(*fwd, fw, xq, reps=3)
    np.testing.assert_array_equal(q3, q1 * np.uint64(3))
    np.testing.assert_array_equal(oracle.spmv_push_fixed(*fwd, fw, xq, threads=4), q1)
    deg = np.maximum(g.in_degree, 1).astype(np.float64)
    # each Q32.32 term is within 2^-32 of its product (and the product of the FP32 path is
    # rounded once), so the two sums differ by rounding only
    diff = np.abs(q1.astype(np.float64) / 2 ** 32 - y1)
    assert np.all(diff <= 1e-5 * np.abs(y1) + 4 * deg * 2.0 ** -32)
