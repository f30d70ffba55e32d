"""GPU parity: the HIP path through the C-ABI against the CPU oracle.

Bars: integer / fixed-point results bit-exact; FP32 ranks within 1e-5 max-abs per vertex of
the oracle's FP32 path (the reference's own FP32 arithmetic), plus two per-row bars in scaled
units that hold heavy and light rows alike (tests/fp32_bars.py): one iteration from the same
start within the error bound of a reordered sum of the row's terms, and after a fixed number
of iterations every row at most twice as far from FP64 as the oracle's row plus 64 ulp.
"""
import json
import os
import sys

import numpy as np
import pytest

import accelgraph_amd as ag
import oracle
from fp32_bars import anchor_check, f32_row_bars, one_step_check

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "rtl_gather_kat.json")))
FP32_TOL = 1e-5

# plan thresholds small enough that a scale-11..14 graph populates every bin, including
# hub rows split over several segments
SMALL = dict(thread_max_deg=2, group16_max_deg=6, wave_max_deg=20, block_max_deg=60,
             tile_rows=7, tile_nnz=9, segment_nnz=256, order=0)
PARAMS = {"default": None, "small": SMALL, "small-order1": dict(SMALL, order=1),
          "tiles": dict(thread_max_deg=16, tile_rows=1024, tile_nnz=4096),
          # source blocking forced on small graphs: 2^10-source blocks, rows of in-degree > 16
          # blocked; every row blocked with 2^8-source blocks; small thresholds (every virtual
          # row class and every combine bin populated); "blocked" also stages up to 128 hot
          # source segments in LDS (the default plan stages none)
          "blocked": dict(src_block_bits=10, blocked_min_bin=1, hot_segments=128),
          "blocked-all": dict(src_block_bits=8, blocked_min_bin=0),
          "blocked-small": dict(SMALL, src_block_bits=9, blocked_min_bin=1),
          # the same blocking without hot segments (every gather through L2)
          "blocked-nohot": dict(src_block_bits=10, blocked_min_bin=1, hot_segments=0),
          # every in-edge its own piece (the first pass a plain gather), tiles by ticket
          "blocked-edge": dict(src_block_bits=9, blocked_min_bin=0, piece_nnz=1,
                               tile_order=ag.TILES_TICKET),
          # the last 4096 sources read from an LDS table in the FP32 first pass, and not
          "blocked-top": dict(src_block_bits=10, blocked_min_bin=0, top_table=ag.TOP_ON),
          "blocked-notop": dict(src_block_bits=10, blocked_min_bin=0, top_table=ag.TOP_OFF),
          # first-pass ids in stream order rather than sorted by source (sorted is the
          # default without hot segments), with multi-edge and with single-edge pieces
          "blocked-unsorted": dict(src_block_bits=10, blocked_min_bin=1, tile_sort=ag.SORT_OFF),
          "blocked-edge-unsorted": dict(src_block_bits=9, blocked_min_bin=0, piece_nnz=1,
                                        tile_sort=ag.SORT_OFF)}
BLOCKED = ["blocked", "blocked-all", "blocked-small", "blocked-nohot", "blocked-edge",
           "blocked-top", "blocked-notop", "blocked-unsorted", "blocked-edge-unsorted"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def upload(device, g, name="default", weights=False):
    p = PARAMS[name]
    return device.upload(g, weights=weights, params=ag.plan_params(**p) if p else None)


def ref_pr(g, **kw):
    return oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src, **kw)


def f64_ref(g):
    return oracle.pagerank_f64(g.out_degree, g.inv_idx, g.in_degree, g.inv_src, tol=1e-15,
                               max_iterations=3000, threads=8)


# ------------------------------------------------------------------ conversions

@pytest.mark.parametrize("bits", [64, 32, 16, 8])
def test_fixed_conversions_bit_exact(device, bits):
    say(f"fixed conversions bits={bits}")
    rng = np.random.default_rng(bits)
    xs = np.concatenate([
        np.array([0.0, -1.0, np.nan, np.inf, 1.0, 0.5, 2.0 ** -17, 2.0 ** -18, 0.15, 0.85,
                  65535.996, 65536.0, 1e30, 1e-30, 255.9, 256.0, 15.97, 16.0, 2.0 ** 31,
                  2.0 ** 32], np.float32),
        rng.random(4000).astype(np.float32) * 8,
        (rng.random(4000) * 2.0 ** rng.integers(-40, 40, 4000)).astype(np.float32)])
    np.testing.assert_array_equal(device.fixed_encode(bits, xs), oracle.fixed_encode(bits, xs))
    qs = np.concatenate([np.array([0, 1, 2, 3, 65535, 65536, 2 ** 24 + 1, 2 ** 53 + 1,
                                   2 ** 63 + 12345, 2 ** 64 - 1], np.uint64),
                         rng.integers(0, 2 ** 63, 4000, dtype=np.uint64),
                         rng.integers(0, 2 ** 30, 4000, dtype=np.uint64)])
    np.testing.assert_array_equal(device.fixed_decode(bits, qs), oracle.fixed_decode(bits, qs))


# ------------------------------------------------------------------ RTL known answers

def _expanded(case, xkey):
    """edge j of the fixture reads its own value: vertex n + j carries it"""
    n, vals = case["n"], case[xkey]
    m = len(vals)
    idx = np.array(case["idx"], np.uint32)
    deg = np.array(case["deg"], np.uint32)
    N = n + m
    full_idx = np.concatenate([idx, np.full(m, m, np.uint32)]) if m else idx
    full_deg = np.concatenate([deg, np.zeros(m, np.uint32)])
    src = (n + np.arange(m)).astype(np.uint32)
    return N, full_idx, full_deg, src


@pytest.mark.parametrize("pname", ["default", "small"])
@pytest.mark.parametrize("case", [c for c in KAT["cases"] if c["op"] == "pagerank_gather"
                                  and c["n"] > 0],
                         ids=lambda c: f"{c['kind']}-{c['fixture']}-{c['cls']}")
def test_gather_rtl_kat(device, case, pname):
    say(f"gather kat {case['kind']} {case['fixture']} {case['cls']} {pname}")
    N, idx, deg, src = _expanded(case, "values")
    dg = device.upload(n=N, out_degree=np.zeros(N, np.uint32), inv_idx=idx, in_degree=deg,
                       inv_src=src, params=ag.plan_params(**PARAMS[pname]) if PARAMS[pname] else None)
    n = case["n"]
    vals = [int(v) for v in case["values"]]
    if case["kind"] == "float":
        full = np.zeros(N, np.uint32)
        full[n:] = vals
        got = dg.gather(full.view(np.float32), ag.GATHER_F32).view(np.uint32)[:n]
        exp = np.array([int(v) for v in case["expected"]], np.uint32)
        if pname == "default":
            np.testing.assert_array_equal(got, exp)  # thread bin: stream order, bit-exact
        else:
            ef = exp.view(np.float32)
            gf = got.view(np.float32)
            np.testing.assert_allclose(gf, ef, rtol=1e-6)
    else:
        gtype = ag.GATHER_U64 if case["kind"] == "fixed64" else ag.GATHER_U32
        full = np.zeros(N, np.uint64 if gtype == ag.GATHER_U64 else np.uint32)
        full[n:] = vals
        got = dg.gather(full, gtype)[:n]
        np.testing.assert_array_equal(got, np.array([int(v) for v in case["expected"]], np.uint64))
    dg.close()


@pytest.mark.parametrize("case", [c for c in KAT["cases"] if c["op"] == "spmv_gather"
                                  and c["n"] > 0],
                         ids=lambda c: f"{c['kind']}-{c['fixture']}-{c['cls']}")
def test_spmv_rtl_kat(device, case):
    say(f"spmv kat {case['kind']} {case['fixture']} {case['cls']}")
    N, idx, deg, src = _expanded(case, "x")
    n, m = case["n"], len(case["x"])
    xs = [int(v) for v in case["x"]]
    ws = [int(v) for v in case["w"]]
    exp = np.array([int(v) for v in case["expected"]])
    if case["kind"] == "float":
        w = np.array(ws, np.uint32).view(np.float32)
        dg = device.upload(n=N, out_degree=np.zeros(N, np.uint32), inv_idx=idx, in_degree=deg,
                           inv_src=src, inv_weight=w)
        x = np.zeros(N, np.uint32)
        x[n:] = xs
        y, _ = dg.spmv(x.view(np.float32))
        np.testing.assert_array_equal(y.view(np.uint32)[:n], exp.astype(np.uint32))
    else:
        dg = device.upload(n=N, out_degree=np.zeros(N, np.uint32), inv_idx=idx, in_degree=deg,
                           inv_src=src)
        import ctypes as C
        u32p = C.POINTER(C.c_uint32)
        wq = np.array(ws, np.uint32)
        st = ag.lib().ag_graph_set_fixed_weights(dg.handle, idx.ctypes.data_as(u32p),
                                                 deg.ctypes.data_as(u32p),
                                                 wq.ctypes.data_as(u32p))
        assert st == 0, ag.lib().ag_last_error()
        x = np.zeros(N, np.uint32)
        x[n:] = xs
        y, _ = dg.spmv_fixed(x)
        np.testing.assert_array_equal(y[:n], exp.astype(np.uint64))
    dg.close()


# ------------------------------------------------------------------ gathers, every bin

@pytest.mark.parametrize("pname", list(PARAMS))
def test_gather_all_bins(device, pname):
    say(f"gather all bins {pname}")
    g = ag.Graph.rmat(13)
    dg = upload(device, g, pname)
    info = dg.info()
    if pname.startswith("small"):
        assert all(info.bin_rows[b] > 0 for b in range(5)) and info.num_segments > info.bin_rows[4]
    rng = np.random.default_rng(7)
    for gtype, vt in [(ag.GATHER_U64, np.uint64), (ag.GATHER_U32, np.uint32),
                      (ag.GATHER_U16, np.uint16), (ag.GATHER_U8, np.uint8)]:
        hi = np.iinfo(vt).max
        v = rng.integers(0, hi, g.n, dtype=vt, endpoint=True)
        exp = oracle.gather_u64(g.inv_idx, g.in_degree, g.inv_src, v.astype(np.uint64))
        np.testing.assert_array_equal(dg.gather(v, gtype), exp)
    v = rng.random(g.n).astype(np.float32)
    exp = oracle.gather_f32(g.inv_idx, g.in_degree, g.inv_src, v)
    got = dg.gather(v, ag.GATHER_F32)
    np.testing.assert_allclose(got, exp, rtol=2e-6, atol=1e-6)
    # thread-bin rows are summed in stream order: bit-exact
    tb = g.in_degree <= info.params.thread_max_deg
    np.testing.assert_array_equal(got[tb], exp[tb])
    # deterministic from run to run
    np.testing.assert_array_equal(dg.gather(v, ag.GATHER_F32), got)
    dg.close()


# ------------------------------------------------------------------ PageRank FP32

@pytest.mark.parametrize("pname", list(PARAMS))
def test_pagerank_f32_reference_rule(device, rmat16, pname):
    say(f"pagerank f32 reference rule {pname}")
    g = rmat16
    dg = upload(device, g, pname)
    info = dg.info()
    if pname == "blocked":    # hot segments gathered from LDS, the rest through L2
        assert info.hot_segments > 0 and 0 < info.hot_edges < g.m
    if pname == "blocked-nohot":
        assert info.hot_segments == 0 and info.src_block_bits == 10
    run = dg.pagerank(max_iterations=100, epsilon=1e-4)
    ref = ref_pr(g, max_iterations=100, epsilon=1e-4)
    # Rows outside the thread bin are summed in a fixed tree order rather than the reference's
    # sequential order. Near the end of the run the deltas of the largest ranks are a few ulps
    # (ulp(244.0f) = 1.5e-5 against epsilon = 1e-4), so the last active vertex can flicker:
    # the active-count rule may stop a couple of iterations apart while every rank stays
    # within the FP32 tolerance.
    assert abs(run.iterations - ref.iterations) <= 3
    k = min(run.executed, ref.executed)
    assert np.all(np.abs(run.hist_active[:k].astype(np.int64) -
                         ref.hist_active[:k].astype(np.int64)) <= 10 + g.n // 1000)
    err = np.abs(run.ranks - ref.ranks).max()
    assert err <= FP32_TOL
    np.testing.assert_allclose(run.ranks, ref.ranks, rtol=5e-5)
    # late deltas sit at the FP32 rounding floor (single-ulp rank changes)
    np.testing.assert_allclose(run.hist_err[:k], ref.hist_err[:k], rtol=1e-3, atol=1e-7)
    assert abs(run.rank_sum - ref.rank_sum) < 1e-6
    f64 = f64_ref(g)
    gpu_err, cpu_err = np.abs(run.ranks - f64).max(), np.abs(ref.ranks - f64).max()
    say(f"  iterations={run.iterations} maxabs_vs_cpu={err:.3e} vs_f64 gpu={gpu_err:.3e} "
        f"cpu={cpu_err:.3e}")
    assert gpu_err <= 2 * cpu_err + 1e-9
    f32_row_bars(dg, g, 30)
    dg.close()


REF_ORDER_PLANS = {"default": {}, "small": SMALL, "small-order1": dict(SMALL, order=1),
                   "tiles": dict(tile_rows=7, tile_nnz=64),
                   "blocked": dict(src_block_bits=10, blocked_min_bin=1)}


@pytest.mark.parametrize("pname", list(REF_ORDER_PLANS))
def test_pagerank_f32_reference_order(device, rmat16, pname):
    """fp32_order = AG_ORDER_REFERENCE: every row summed in sequence in inverse-CSR order (the
    reference's CPU loop, pageRank.c:2903-2907, and float sum kernel,
    pagerank_kernel_tb.sv:303-314) -- rows that fit a tile by one lane from LDS, longer rows
    (here up to 6310 in-edges: two LDS chunks) by one lane over chunks. Under the reference's
    active-count stop (pageRank.c:1065-1074) the run equals the sequential oracle bit for bit:
    iteration count, active history, scaled and normalised ranks, whatever the other plan
    parameters (source blocking is turned off by the mode)."""
    say(f"pagerank f32 reference order {pname}")
    g = rmat16
    assert g.in_degree.max() > 4096
    dg = device.upload(g, params=ag.plan_params(**REF_ORDER_PLANS[pname],
                                                 fp32_order=ag.ORDER_REFERENCE))
    info = dg.info()
    assert info.src_block_bits == 0
    assert all(info.bin_rows[b] == 0 for b in (1, 2, 3))   # tile rows and whole-row hubs only
    run = dg.pagerank(max_iterations=100, epsilon=1e-4, scaled=True)
    ref = ref_pr(g, max_iterations=100, epsilon=1e-4)
    assert run.iterations == ref.iterations and run.executed == ref.executed
    np.testing.assert_array_equal(run.hist_active, ref.hist_active)
    np.testing.assert_array_equal(run.scaled.view(np.uint32), ref.scaled.view(np.uint32))
    np.testing.assert_array_equal(run.ranks.view(np.uint32), ref.ranks.view(np.uint32))
    np.testing.assert_allclose(run.hist_err, ref.hist_err, rtol=1e-9, atol=1e-15)
    # the uniform start and the L1 rule too
    run = dg.pagerank(max_iterations=100, stop_rule=ag.STOP_L1, l1_tolerance=1e-6,
                      init=ag.INIT_UNIFORM)
    ref = ref_pr(g, max_iterations=100, stop_rule=1, l1_tolerance=1e-6, init=1)
    assert run.iterations == ref.iterations
    np.testing.assert_array_equal(run.ranks.view(np.uint32), ref.ranks.view(np.uint32))
    say(f"  iterations={ref.iterations} bit-exact")
    dg.close()


def test_pagerank_f32_config1_l1_rule(device, rmat16):
    """RMAT-16, damping 0.85, uniform 1/N start, stop at L1 <= 1e-4 or 100 iterations."""
    say("pagerank config-1 l1 rule")
    g = rmat16
    dg = device.upload(g)
    run = dg.pagerank(max_iterations=100, stop_rule=ag.STOP_L1, l1_tolerance=1e-4,
                      init=ag.INIT_UNIFORM)
    ref = ref_pr(g, max_iterations=100, stop_rule=1, l1_tolerance=1e-4, init=1)
    assert run.iterations == ref.iterations and run.error_total <= 1e-4
    assert np.abs(run.ranks - ref.ranks).max() <= FP32_TOL
    f64 = oracle.pagerank_f64(g.out_degree, g.inv_idx, g.in_degree, g.inv_src, tol=0.0,
                              max_iterations=run.executed, threads=8)
    anchor_check(run.ranks, ref.ranks, f64)
    f32_row_bars(dg, g, run.executed)
    dg.close()


def test_pagerank_f32_user_init_and_determinism(device, rmat12):
    say("pagerank user init")
    g = rmat12
    dg = upload(device, g, "small")
    rng = np.random.default_rng(5)
    init = (rng.random(g.n) * 2).astype(np.float32)
    run = dg.pagerank(max_iterations=30, stop_rule=ag.STOP_NONE, init_ranks=init)
    ref = ref_pr(g, max_iterations=30, stop_rule=2, init=2, init_ranks=init)
    assert run.iterations == 30 == ref.iterations
    assert np.abs(run.ranks - ref.ranks).max() <= FP32_TOL
    run2 = dg.pagerank(max_iterations=30, stop_rule=ag.STOP_NONE, init_ranks=init)
    np.testing.assert_array_equal(run.ranks, run2.ranks)
    dg.close()


# ------------------------------------------------------------------ fixed point

@pytest.mark.parametrize("bits", [64, 32, 16, 8])
@pytest.mark.parametrize("pname", ["default", "small"] + BLOCKED)
def test_pagerank_fixed_bit_exact(device, rmat16, bits, pname):
    say(f"pagerank fixed{bits} {pname}")
    g = rmat16
    dg = upload(device, g, pname)
    run = dg.pagerank(kind=bits, max_iterations=60, epsilon=1e-4, scaled=True)
    ref = ref_pr(g, kind=bits, max_iterations=60, epsilon=1e-4)
    assert run.iterations == ref.iterations and run.executed == ref.executed
    np.testing.assert_array_equal(run.hist_active, ref.hist_active)
    np.testing.assert_array_equal(run.scaled.view(np.uint32), ref.scaled.view(np.uint32))
    np.testing.assert_array_equal(run.ranks.view(np.uint32), ref.ranks.view(np.uint32))
    np.testing.assert_allclose(run.hist_err, ref.hist_err, rtol=1e-9, atol=1e-15)
    dg.close()


QUANT = [ag.KIND_QUANT32, ag.KIND_QUANT16, ag.KIND_QUANT8]


@pytest.mark.parametrize("kind", QUANT)
@pytest.mark.parametrize("pname", ["default", "small"] + BLOCKED)
def test_pagerank_quant_bit_exact(device, rmat16, kind, pname):
    """Quantised pulls: exact u64 code sums, so every plan gives the oracle's ranks bit for bit
    (iteration count, per-iteration active counts and ranks)."""
    say(f"pagerank quant{kind - 100} {pname}")
    g = rmat16
    dg = upload(device, g, pname)
    run = dg.pagerank(kind=kind, max_iterations=60, epsilon=1e-4, scaled=True)
    ref = ref_pr(g, kind=kind, max_iterations=60, epsilon=1e-4)
    assert run.iterations == ref.iterations and run.executed == ref.executed
    np.testing.assert_array_equal(run.hist_active, ref.hist_active)
    np.testing.assert_array_equal(run.scaled.view(np.uint32), ref.scaled.view(np.uint32))
    np.testing.assert_array_equal(run.ranks.view(np.uint32), ref.ranks.view(np.uint32))
    np.testing.assert_allclose(run.hist_err, ref.hist_err, rtol=1e-9, atol=1e-15)
    dg.close()


@pytest.mark.parametrize("kind", QUANT)
def test_pagerank_quant_inits_and_corners(device, rmat12, kind):
    """User start vectors with zero, negative and huge entries (codes 0 and saturation), the
    uniform start with the L1 stop rule, and graphs whose contributions are all zero (scale 1)."""
    say(f"pagerank quant{kind - 100} inits")
    g = rmat12
    rng = np.random.default_rng(kind)
    init = rng.random(g.n).astype(np.float32) * 3
    init[::7] = 0.0
    init[1::11] = -2.5
    init[5] = 1e30
    dg = device.upload(g)
    for kw in (dict(init_ranks=init, max_iterations=12, stop_rule=ag.STOP_NONE),
               dict(init=ag.INIT_UNIFORM, stop_rule=ag.STOP_L1, l1_tolerance=1e-6,
                    max_iterations=200)):
        run = dg.pagerank(kind=kind, scaled=True, **kw)
        okw = dict(kw)
        if "init_ranks" in okw:
            okw["init"] = 2
        ref = ref_pr(g, kind=kind, **okw)
        assert run.iterations == ref.iterations
        np.testing.assert_array_equal(run.scaled.view(np.uint32), ref.scaled.view(np.uint32))
    dg.close()
    for src, dst, n in [([], [], 5), ([0, 1], [1, 0], 2), ([1] * 3, [0] * 3, 4)]:
        h = ag.Graph.from_edges(src, dst, n=n, simplify=False)
        dh = device.upload(h)
        run = dh.pagerank(kind=kind, max_iterations=50)
        ref = ref_pr(h, kind=kind, max_iterations=50)
        assert run.iterations == ref.iterations
        np.testing.assert_array_equal(run.ranks.view(np.uint32), ref.ranks.view(np.uint32))
        dh.close()


# ------------------------------------------------------------------ push / data-driven

@pytest.mark.parametrize("kind", [64, 32, 16, 8])
@pytest.mark.parametrize("pname", ["default", "blocked-all"])
def test_pagerank_push_fixed_bit_exact(device, rmat16, kind, pname):
    """Push direction, fixed point: the binned u64 sums are exact, so the ranks equal
    the pull oracle's bit for bit (the plan only relabels rows: the forward CSR is built in
    device row ids for each)."""
    say(f"pagerank push fixed{kind} {pname}")
    g = rmat16
    dg = upload(device, g, pname)
    run = dg.pagerank(kind=kind, direction=ag.DIR_PUSH, max_iterations=60, epsilon=1e-4,
                      scaled=True)
    ref = ref_pr(g, kind=kind, max_iterations=60, epsilon=1e-4)
    assert run.iterations == ref.iterations and run.executed == ref.executed
    np.testing.assert_array_equal(run.hist_active, ref.hist_active)
    np.testing.assert_array_equal(run.ranks.view(np.uint32), ref.ranks.view(np.uint32))
    dg.close()


def _fp32_bar(run, ref, iteration_slack=3):
    """iteration_slack None: the stopping iteration is only required to fall in the oracle's
    tail (at most 0.1% of the vertices still active there)."""
    if iteration_slack is None:
        k = min(run.iterations, ref.executed - 1)
        assert ref.hist_active[k] <= max(1, ref.ranks.size // 1000)
    else:
        assert abs(run.iterations - ref.iterations) <= iteration_slack
    assert np.abs(run.ranks - ref.ranks).max() <= FP32_TOL
    np.testing.assert_allclose(run.ranks, ref.ranks, rtol=5e-5)


@pytest.mark.parametrize("pname", ["default", "small", "blocked"])
def test_pagerank_push_f32(device, rmat16, pname):
    """Push direction, FP32: the LDS atomics add in an order of their own, different on every
    run, so the bar is the FP32 one. With a fixed iteration count the ranks are compared
    directly. Under the reference's stop rule the last vertices at the epsilon threshold have
    ulp-sized deltas, so how long that tail lasts is rounding noise: the run must stop inside
    the oracle's tail, with ranks within the FP32 bar."""
    say(f"pagerank push f32 {pname}")
    g = rmat16
    dg = upload(device, g, pname)
    run = dg.pagerank(direction=ag.DIR_PUSH, max_iterations=40, stop_rule=ag.STOP_NONE)
    ref = ref_pr(g, max_iterations=40, stop_rule=2)
    _fp32_bar(run, ref, iteration_slack=0)
    run = dg.pagerank(direction=ag.DIR_PUSH, max_iterations=100, epsilon=1e-4)
    ref = ref_pr(g, max_iterations=100, epsilon=1e-4)
    _fp32_bar(run, ref, iteration_slack=None)
    f64 = f64_ref(g)
    assert np.abs(run.ranks - f64).max() <= 2 * np.abs(ref.ranks - f64).max() + 1e-9
    dg.close()


def assert_eps_fixed_point(g, scaled, eps):
    """One more (sequential) evaluation of every vertex moves it by less than epsilon, up to
    the rounding of a sum taken in another order: a d-term sum of positive terms differs
    between orders by at most about d * 2^-24 * sum, and sum < rank / 0.85."""
    one = oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                               max_iterations=1, stop_rule=2, init=2, init_ranks=scaled)
    order = (g.in_degree.astype(np.float64) + 2) * 2.0 ** -24 * np.abs(scaled) / 0.85 * 2
    assert np.all(np.abs(one.scaled.astype(np.float64) - scaled) < eps + order)


def dd_ref(g, **kw):
    return oracle.pagerank_dd_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                                   g.edges_idx, g.out_degree, g.fwd_dest, **kw)


@pytest.mark.parametrize("pname", ["default", "small", "blocked", "blocked-all",
                                   "blocked-small"])
def test_pagerank_data_driven(device, rmat16, pname):
    """Data-driven pull against the oracle's. On unblocked plans it keeps worklists, rows keep
    their edge order and rows up to in-degree 8192 (all of RMAT-16's) are summed in that order,
    so the run is the oracle's bit for bit. On a source-blocked plan it runs as a full pull with
    sign-bit lists, summed in the FP32 pull's order, so it gets the FP32 bar there, and with
    epsilon 0 it equals the FP32 pull on the same plan bit for bit. Either way the stopping
    state satisfies the invariant: one more evaluation of any vertex moves it by less than
    epsilon."""
    say(f"pagerank data-driven {pname}")
    g = rmat16
    assert g.in_degree.max() <= 8192
    dg = upload(device, g, pname)
    eps = 1e-4
    run = dg.pagerank(direction=ag.DIR_DATA_DRIVEN, max_iterations=300, epsilon=eps,
                      scaled=True)
    ref = dd_ref(g, max_iterations=300, epsilon=eps)
    if pname in BLOCKED:
        assert abs(run.iterations - ref.iterations) <= 10
        assert np.abs(run.ranks - ref.ranks).max() <= FP32_TOL
    else:
        assert run.iterations == ref.iterations and run.executed == ref.executed
        np.testing.assert_array_equal(run.hist_active, ref.hist_active)
        np.testing.assert_array_equal(run.scaled.view(np.uint32), ref.scaled.view(np.uint32))
        np.testing.assert_allclose(run.hist_err, ref.hist_err, rtol=1e-9, atol=1e-15)
    assert run.executed == run.iterations + 1 and run.hist_active[-1] == 0
    assert_eps_fixed_point(g, run.scaled, eps)
    # epsilon 0: every listed vertex updates, and the ranks are the plain pull's
    run = dg.pagerank(direction=ag.DIR_DATA_DRIVEN, max_iterations=12, epsilon=0.0)
    ref = ref_pr(g, max_iterations=12, stop_rule=2)
    assert np.abs(run.ranks - ref.ranks).max() <= FP32_TOL
    np.testing.assert_allclose(run.ranks, ref.ranks, rtol=5e-5)
    if pname in BLOCKED:
        pull = dg.pagerank(max_iterations=12, stop_rule=ag.STOP_NONE)
        assert dg.info().src_block_bits > 0
        np.testing.assert_array_equal(run.ranks.view(np.uint32), pull.ranks.view(np.uint32))
    dg.close()


def test_pagerank_data_driven_hub_rows(device):
    """Rows of in-degree above 8192 are listed apart and gathered by a whole workgroup in a
    fixed tree order: the FP32 bar and the stopping invariant, on a graph with a row of
    in-degree 12000 and a vertex whose 9000 out-edges all get marked when it changes."""
    say("pagerank data-driven hub rows")
    rng = np.random.default_rng(5)
    n = 30000
    src = [np.arange(1, 12001), np.full(9000, 7), rng.integers(0, n, 60000)]
    dst = [np.zeros(12000, np.int64), np.arange(20000, 29000), rng.integers(0, n, 60000)]
    src.append(np.arange(100, 200))
    dst.append(np.arange(101, 201))
    g = ag.Graph.from_edges(np.concatenate(src), np.concatenate(dst), n=n)
    assert g.in_degree.max() > 8192 and g.out_degree.max() > 8192
    eps = 1e-5
    for pname in ("default", "blocked"):
        dg = upload(device, g, pname)
        run = dg.pagerank(direction=ag.DIR_DATA_DRIVEN, max_iterations=300, epsilon=eps,
                          scaled=True)
        ref = dd_ref(g, max_iterations=300, epsilon=eps)
        assert abs(run.iterations - ref.iterations) <= 10
        assert np.abs(run.ranks - ref.ranks).max() <= FP32_TOL
        assert_eps_fixed_point(g, run.scaled, eps)
        dg.close()


def test_pagerank_data_driven_user_starts(device, rmat16):
    """On a source-blocked plan a non-negative start runs as the full pull with sign-bit lists;
    a start with negative entries or -0 keeps explicit worklists (the sign bit is not free).
    Both against the oracle's data-driven pull (FP32 bar) and the stopping invariant."""
    say("pagerank data-driven user starts")
    g = rmat16
    dg = upload(device, g, "blocked")
    rng = np.random.default_rng(9)
    pos = (rng.random(g.n) * 2).astype(np.float32)
    signed = pos.copy()
    signed[::5] *= -1
    signed[7] = -0.0
    for init in (pos, signed):
        # three iterations still depend on the start: a sign taken for a mark would move
        # ranks by O(1); block-ordered sums of mixed-sign terms differ by far less than 1e-3
        run = dg.pagerank(direction=ag.DIR_DATA_DRIVEN, init_ranks=init, max_iterations=3,
                          epsilon=1e-4, stop_rule=ag.STOP_NONE, scaled=True)
        ref = dd_ref(g, init=2, init_ranks=init, max_iterations=3, epsilon=1e-4)
        np.testing.assert_allclose(run.scaled, ref.scaled, rtol=1e-3, atol=1e-4)
        assert np.all(np.abs(run.hist_active.astype(np.int64) -
                             ref.hist_active.astype(np.int64)) <= 10 + g.n // 1000)
        run = dg.pagerank(direction=ag.DIR_DATA_DRIVEN, init_ranks=init, max_iterations=300,
                          epsilon=1e-4, scaled=True)
        ref = dd_ref(g, init=2, init_ranks=init, max_iterations=300, epsilon=1e-4)
        assert abs(run.iterations - ref.iterations) <= 10
        assert np.abs(run.ranks - ref.ranks).max() <= FP32_TOL
        assert_eps_fixed_point(g, run.scaled, 1e-4)
    dg.close()


def test_push_bin_shapes(device):
    """The push's bins (ranges of at most 4096 destination rows and 2^17 in-edges; a heavier
    row is a bin by itself) on a graph that reaches every case: two hub rows above 2^17
    in-edges, a run of rows without in-edges, a 240 K-row chain (bins cut at 4096 rows) and a
    vertex count that is not a multiple of 4096. Fixed point bit-exact with the pull oracle,
    FP32 within the FP32 bar; SpMV push likewise."""
    say("push bins: hub rows, empty rows, long chain")
    n = 250_003
    hub0 = np.arange(1, 200_001, dtype=np.uint32)
    hub1 = np.arange(60_000, 200_001, dtype=np.uint32)
    chain = np.arange(10_000, n - 1, dtype=np.uint32)
    src = np.concatenate([hub0, hub1, chain, [n - 1]]).astype(np.uint32)
    dst = np.concatenate([np.zeros_like(hub0), np.ones_like(hub1), chain + 1, [5]])
    g = ag.Graph.from_edges(src, dst.astype(np.uint32), n=n)
    assert g.in_degree[0] > 131072 and g.in_degree[1] > 131072
    dg = device.upload(g)
    for kind in (64, 32):
        run = dg.pagerank(kind=kind, direction=ag.DIR_PUSH, max_iterations=30, scaled=True)
        ref = ref_pr(g, kind=kind, max_iterations=30)
        assert run.iterations == ref.iterations
        np.testing.assert_array_equal(run.ranks.view(np.uint32), ref.ranks.view(np.uint32))
    run = dg.pagerank(direction=ag.DIR_PUSH, max_iterations=20, stop_rule=ag.STOP_NONE)
    ref = ref_pr(g, max_iterations=20, stop_rule=2)
    assert run.iterations == ref.iterations == 20
    # A 200 K-term FP32 sum in sequence (the oracle) is off by about n * 2^-24 / 4 relative
    # (3e-3 here); the push sums a hub row in a tree. So the FP32 bar holds on the ordinary
    # rows, and every row must be as close to the same 20 iterations in FP64 as the oracle is.
    light = g.in_degree <= 4096
    assert np.abs(run.ranks - ref.ranks)[light].max() <= FP32_TOL
    np.testing.assert_allclose(run.ranks[light], ref.ranks[light], rtol=5e-5)
    esrc = g.inv_src.astype(np.int64)
    erow = np.repeat(np.arange(n), g.in_degree.astype(np.int64))
    od = g.out_degree.astype(np.float64)
    r = np.full(n, 0.15)
    for _ in range(20):
        c = np.where(od > 0, r / np.maximum(od, 1), 0.0)
        r = 0.15 + 0.85 * np.bincount(erow, weights=c[esrc], minlength=n)
    r /= n
    assert np.abs(run.ranks - r).max() <= 2 * np.abs(ref.ranks - r).max() + 1e-9
    rng = np.random.default_rng(5)
    x = rng.random(n).astype(np.float32)
    fwd = (g.edges_idx, g.out_degree, g.fwd_dest)
    y, _ = dg.spmv_push(x, 2)
    ref = oracle.spmv_push_f32(*fwd, None, x, 2)
    bound = (2 * g.in_degree.astype(np.float64) + 2) * 2.0 ** -24 * 2 * np.abs(ref)
    assert np.all(np.abs(y.astype(np.float64) - ref) <= bound + 1e-30)
    xq = oracle.fixed_encode(64, x)
    yq, _ = dg.spmv_push_fixed(xq, 2)
    np.testing.assert_array_equal(yq, oracle.spmv_push_fixed(*fwd, None, xq, 2))
    dg.close()


def test_pagerank_directions_corners(device):
    say("push / data-driven corner cases")
    for src, dst, n in [([], [], 1), ([], [], 6), ([0], [1], 2), ([0, 1, 2], [1, 2, 0], 3),
                        ([1] * 5, [0] * 5, 7), ([0, 0, 0, 0], [1, 2, 3, 4], 5)]:
        g = ag.Graph.from_edges(src, dst, n=n, simplify=False)
        dg = device.upload(g)
        run = dg.pagerank(direction=ag.DIR_PUSH, kind=32, max_iterations=100)
        ref = ref_pr(g, kind=32, max_iterations=100)
        assert run.iterations == ref.iterations
        np.testing.assert_array_equal(run.ranks.view(np.uint32), ref.ranks.view(np.uint32))
        run = dg.pagerank(direction=ag.DIR_PUSH, max_iterations=100)
        ref = ref_pr(g, max_iterations=100)
        assert run.iterations == ref.iterations
        np.testing.assert_allclose(run.ranks, ref.ranks, rtol=1e-6)
        run = dg.pagerank(direction=ag.DIR_DATA_DRIVEN, max_iterations=100)
        ref = dd_ref(g, max_iterations=100)
        assert run.iterations == ref.iterations
        np.testing.assert_allclose(run.ranks, ref.ranks, rtol=1e-6)
        dg.close()
    g = ag.Graph.rmat(8)
    dg = device.upload(g)
    for kw in (dict(direction=ag.DIR_PUSH, kind=ag.KIND_QUANT16),
               dict(direction=ag.DIR_PUSH, storage=ag.STORE_BF16),
               dict(direction=ag.DIR_DATA_DRIVEN, kind=32),
               dict(direction=7)):
        with pytest.raises(ag.AGError):
            dg.pagerank(max_iterations=3, **kw)
    dg.close()


@pytest.mark.parametrize("scale", [9, 12])
def test_pagerank_lds_table_whole_vector(device, scale):
    """A graph of at most 4096 vertices: the first pass's LDS table holds every source, so
    every gather reads LDS. FP32 ranks equal the table-less plan's bit for bit (also with
    16-bit storage), fixed-32 equals the oracle."""
    say(f"lds table holding the whole vector, rmat-{scale}")
    g = ag.Graph.rmat(scale)
    base = dict(src_block_bits=8, blocked_min_bin=0)
    on = device.upload(g, params=ag.plan_params(**base, top_table=ag.TOP_ON))
    off = device.upload(g, params=ag.plan_params(**base, top_table=ag.TOP_OFF))
    assert on.info().top_sources == g.n and off.info().top_sources == 0
    kw = dict(max_iterations=15, stop_rule=ag.STOP_NONE)
    for extra in (dict(), dict(storage=ag.STORE_BF16, refine_l1=1e-5)):
        a, b = on.pagerank(**kw, **extra), off.pagerank(**kw, **extra)
        np.testing.assert_array_equal(a.ranks.view(np.uint32), b.ranks.view(np.uint32))
    fx = on.pagerank(kind=32, **kw)
    ref = ref_pr(g, kind=32, max_iterations=15, stop_rule=2)
    np.testing.assert_array_equal(fx.ranks.view(np.uint32), ref.ranks.view(np.uint32))
    on.close()
    off.close()


def test_pagerank_blocked_auto_rmat22(device):
    """Above 2^21 vertices the default plan blocks sources (2^19 per block). RMAT-22: fixed32
    bit-exact and FP32 within tolerance of the oracle for 20 iterations; the blocked, unblocked
    and hot-segment (LDS-staged) FP32 runs all within the per-row FP32 bars."""
    say("pagerank blocked auto rmat-22")
    g = ag.Graph.rmat(22)
    dg = device.upload(g)
    info = dg.info()
    assert info.src_block_bits == 19 and info.virtual_rows > 0
    assert info.hot_segments == 0 and info.src_blocks == 8
    kw = dict(max_iterations=20, stop_rule=ag.STOP_NONE)
    fx = dg.pagerank(kind=32, **kw)
    ref = ref_pr(g, kind=32, max_iterations=20, stop_rule=2, threads=8)
    np.testing.assert_array_equal(fx.ranks.view(np.uint32), ref.ranks.view(np.uint32))
    run = dg.pagerank(**kw)
    ref = ref_pr(g, max_iterations=20, stop_rule=2, threads=8)
    err = np.abs(run.ranks - ref.ranks).max()
    say(f"  fixed32 bit-exact; f32 maxabs={err:.3e}")
    assert err <= FP32_TOL
    f32_row_bars(dg, g, 20)
    dg.close()
    # pass-1 tiles by ticket instead of round-robin, without the LDS table of the last
    # sources (on by default here), and with their ids in stream order instead of sorted by
    # source (the default): the same ranks bit for bit
    assert info.top_sources == 4096 and info.top_edges * 50 >= g.m
    assert info.sorted_tiles == 1
    for prm in (dict(tile_order=ag.TILES_TICKET), dict(top_table=ag.TOP_OFF),
                dict(tile_sort=ag.SORT_OFF)):
        other = device.upload(g, params=ag.plan_params(**prm))
        np.testing.assert_array_equal(other.pagerank(**kw).ranks.view(np.uint32),
                                      run.ranks.view(np.uint32))
        if "top_table" in prm:
            assert other.info().top_sources == 0
        if "tile_sort" in prm:
            assert other.info().sorted_tiles == 0
            np.testing.assert_array_equal(other.pagerank(kind=32, **kw).ranks, fx.ranks)
        other.close()
    # 2^20-source blocks fill the packed word (20 bits of source offset, 12 of position):
    # sorted and unsorted tiles give the same ranks bit for bit
    wide = [device.upload(g, params=ag.plan_params(src_block_bits=20, tile_sort=s))
            for s in (ag.SORT_ON, ag.SORT_OFF)]
    assert [w.info().sorted_tiles for w in wide] == [1, 0]
    assert all(w.info().src_block_bits == 20 for w in wide)
    r_on, r_off = (w.pagerank(**kw).ranks for w in wide)
    np.testing.assert_array_equal(r_on.view(np.uint32), r_off.view(np.uint32))
    f_on, f_off = (w.pagerank(kind=32, **kw).ranks for w in wide)
    np.testing.assert_array_equal(f_on, fx.ranks)
    np.testing.assert_array_equal(f_off, fx.ranks)
    assert np.abs(r_on - run.ranks).max() <= FP32_TOL
    for w in wide:
        w.close()
    plain = device.upload(g, params=ag.plan_params(src_block_bits=ag.SRC_BLOCK_OFF))
    assert plain.info().src_block_bits == 0
    run2 = plain.pagerank(**kw)
    assert np.abs(run.ranks - run2.ranks).max() <= FP32_TOL
    f32_row_bars(plain, g, 20)
    plain.close()
    hot = device.upload(g, params=ag.plan_params(hot_segments=128))
    hinfo = hot.info()
    assert hinfo.hot_segments > 0 and hinfo.src_blocks == 8 + hinfo.hot_segments
    assert hinfo.hot_segment_bits == 14 and 0 < hinfo.hot_edges < g.m
    fx3 = hot.pagerank(kind=32, **kw)
    np.testing.assert_array_equal(fx3.ranks.view(np.uint32), fx.ranks.view(np.uint32))
    run3 = hot.pagerank(**kw)
    assert np.abs(run.ranks - run3.ranks).max() <= FP32_TOL
    f32_row_bars(hot, g, 20)
    hot.close()


def test_rmat25_full_size(device):
    """The benchmark's own graph and plan (RMAT-25, 33.5 M vertices, 528.7 M edges, built on
    the GPU; 2^19-source blocks): the fixed-32 pull and the fixed-64 push bit-exact with the
    oracle after 4 iterations, the FP32 pull within the FP32 bar, and the data-driven pull
    with epsilon 0 equal to the FP32 pull bit for bit."""
    say("rmat-25 full size: build")
    g = ag.Graph.rmat_gpu(device, 25)
    assert g.n == 1 << 25 and g.m == 528_693_006
    dg = device.upload(g)
    info = dg.info()
    assert info.src_block_bits == 19 and info.src_blocks == 64 and info.hot_segments == 0
    kw = dict(max_iterations=4, stop_rule=ag.STOP_NONE)
    say("rmat-25 full size: fixed-32 pull")
    fx = dg.pagerank(kind=32, **kw)
    ref = ref_pr(g, kind=32, max_iterations=4, stop_rule=2, threads=16)
    np.testing.assert_array_equal(fx.ranks.view(np.uint32), ref.ranks.view(np.uint32))
    say("rmat-25 full size: fixed-64 push")
    px = dg.pagerank(kind=64, direction=ag.DIR_PUSH, **kw)
    ref = ref_pr(g, kind=64, max_iterations=4, stop_rule=2, threads=16)
    np.testing.assert_array_equal(px.ranks.view(np.uint32), ref.ranks.view(np.uint32))
    say("rmat-25 full size: FP32 pull and data-driven pull")
    run = dg.pagerank(**kw)
    ref = ref_pr(g, max_iterations=4, stop_rule=2, threads=16)
    assert np.abs(run.ranks - ref.ranks).max() <= FP32_TOL
    light = g.in_degree <= 1024   # rows whose sequential FP32 sum stays within ~1e-4 relative
    np.testing.assert_allclose(run.ranks[light], ref.ranks[light], rtol=1e-4, atol=0)
    say("rmat-25 full size: per-row FP32 bars (hub rows included)")
    f32_row_bars(dg, g, 4, threads=16)
    dd = dg.pagerank(direction=ag.DIR_DATA_DRIVEN, epsilon=0.0, **kw)
    np.testing.assert_array_equal(dd.ranks.view(np.uint32), run.ranks.view(np.uint32))
    dg.close()
    g.close()


# ------------------------------------------------------------------ SpMV

@pytest.mark.parametrize("pname", ["default", "small"] + BLOCKED)
def test_spmv_f32_and_fixed(device, pname):
    say(f"spmv rmat-14 random weights {pname}")
    g = ag.Graph.rmat(14)
    rng = np.random.default_rng(11)
    fw = ((rng.integers(0, 2 ** 24, g.m) + 1) * 2.0 ** -24).astype(np.float32)  # (0, 1]
    g.set_weights(fw)
    x = (rng.integers(0, 2 ** 24, g.n) * 2.0 ** -24).astype(np.float32)  # [0, 1)
    dg = upload(device, g, pname, weights=True)
    info = dg.info()
    assert (info.src_block_bits != 0) == (pname in BLOCKED)
    y, _ = dg.spmv(x, reps=3)
    exp = oracle.spmv_f32(g.inv_idx, g.in_degree, g.inv_src, g.inv_weight, x)
    np.testing.assert_allclose(y, exp, rtol=2e-6, atol=1e-6)
    if info.src_block_bits == 0 or info.params.blocked_min_bin > 0:
        # thread-bin rows that are not source-blocked keep the stream-order sum
        tb = g.in_degree <= info.params.thread_max_deg
        np.testing.assert_array_equal(y[tb], exp[tb])
    xq = oracle.fixed_encode(32, x).astype(np.uint32)
    wq = oracle.fixed_encode(32, g.inv_weight).astype(np.uint32)
    want = oracle.spmv_fixed(g.inv_idx, g.in_degree, g.inv_src, wq, 0, xq)
    yq, _ = dg.spmv_fixed(xq, reps=2)
    np.testing.assert_array_equal(yq, want)
    # caller-supplied weight codes (regular order): same result through the fallback schedule
    import ctypes as C
    u32p = C.POINTER(C.c_uint32)
    idx, deg = np.ascontiguousarray(g.inv_idx, np.uint32), np.ascontiguousarray(g.in_degree, np.uint32)
    wq2 = np.ascontiguousarray(wq[::-1])
    assert ag.lib().ag_graph_set_fixed_weights(dg.handle, idx.ctypes.data_as(u32p),
                                               deg.ctypes.data_as(u32p),
                                               wq2.ctypes.data_as(u32p)) == 0
    yq2, _ = dg.spmv_fixed(xq)
    np.testing.assert_array_equal(yq2, oracle.spmv_fixed(g.inv_idx, g.in_degree, g.inv_src,
                                                         wq2, 0, xq))
    dg.close()


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("pname", ["default", "blocked"])
def test_spmv_push(device, pname, weighted):
    """SpMV push: repetitions accumulate; FP32 within the rounding of reps*d additions taken
    in an order of the kernel's, Q32.32 fixed point bit-exact with the oracle."""
    say(f"spmv push {pname} weighted={weighted}")
    g = ag.Graph.rmat(14)
    rng = np.random.default_rng(21)
    if weighted:
        g.set_weights(rng.random(g.m).astype(np.float32))
    fw = g.fwd_weight if weighted else None
    fwd = (g.edges_idx, g.out_degree, g.fwd_dest)
    dg = upload(device, g, pname, weights=weighted)
    x = rng.random(g.n).astype(np.float32)
    xq = oracle.fixed_encode(64, x)
    for reps in (0, 1, 3):
        y, _ = dg.spmv_push(x, reps)
        ref = oracle.spmv_push_f32(*fwd, fw, x, reps)
        bound = (reps * g.in_degree.astype(np.float64) + 2) * 2.0 ** -24 * 2 * np.abs(ref)
        assert np.all(np.abs(y.astype(np.float64) - ref) <= bound + 1e-30)
        yq, _ = dg.spmv_push_fixed(xq, reps)
        np.testing.assert_array_equal(yq, oracle.spmv_push_fixed(*fwd, fw, xq, reps))
    dg.close()


def test_spmv_unweighted(device):
    say("spmv rmat-14 unweighted")
    g = ag.Graph.rmat(14)
    rng = np.random.default_rng(12)
    x = (rng.integers(0, 2 ** 24, g.n) * 2.0 ** -24).astype(np.float32)  # [0, 1)
    # unweighted graph: constant weight 0.0001f
    for pname in ("default", "blocked-all"):
        dg = upload(device, g, pname)
        y, _ = dg.spmv(x)
        exp = oracle.spmv_f32(g.inv_idx, g.in_degree, g.inv_src, None, x)
        np.testing.assert_allclose(y, exp, rtol=2e-6, atol=1e-9)
        xq = oracle.fixed_encode(32, x).astype(np.uint32)
        yq, _ = dg.spmv_fixed(xq)
        wc = int(oracle.fixed_encode(32, [0.0001])[0])
        np.testing.assert_array_equal(yq, oracle.spmv_fixed(g.inv_idx, g.in_degree, g.inv_src,
                                                            None, wc, xq))
        dg.close()


# ------------------------------------------------------------------ edge cases

def test_watchdog_timeout_then_recovery(device):
    """The bounded wait (the role of waitAFU's run deadline, capienv.c:526-638): with a 1 ms
    deadline, a poll behind 100 queued RMAT-20 iterations (~10 ms of kernels) returns
    AG_ERR_TIMEOUT from the host; the kernels themselves run to their normal end. Without the
    deadline the same session then completes, its fixed-32 ranks bit-exact with the oracle,
    and the device takes new work."""
    say("watchdog: timeout, then the same session completes")
    g = ag.Graph.rmat(20)
    dg = device.upload(g)
    iters = 100
    sess = dg.session(kind=32, max_iterations=iters, stop_rule=ag.STOP_NONE)
    try:
        device.set_watchdog(1, 20)
        sess.step(iters)
        with pytest.raises(ag.AGError) as exc:
            sess.poll()
        assert exc.value.status == ag._abi.AG_ERR_TIMEOUT
    finally:
        device.set_watchdog(0)
    device.sync()
    r = sess.poll()
    assert r.executed == iters
    ranks, _ = sess.end()
    ref = ref_pr(g, kind=32, max_iterations=iters, stop_rule=2, threads=8)
    np.testing.assert_array_equal(ranks.view(np.uint32), ref.ranks.view(np.uint32))
    run = dg.pagerank(kind=32, max_iterations=5, stop_rule=ag.STOP_NONE)
    ref = ref_pr(g, kind=32, max_iterations=5, stop_rule=2, threads=8)
    np.testing.assert_array_equal(run.ranks.view(np.uint32), ref.ranks.view(np.uint32))
    dg.close()


def test_edge_cases(device):
    say("edge cases")
    for src, dst, n in [([], [], 1), ([0], [1], 2), ([0, 1, 2], [1, 2, 0], 3),
                        ([1] * 5, [0] * 5, 7)]:
        g = ag.Graph.from_edges(src, dst, n=n)
        dg = device.upload(g)
        run = dg.pagerank(max_iterations=100)
        ref = ref_pr(g, max_iterations=100)
        assert run.iterations == ref.iterations
        np.testing.assert_array_equal(run.ranks, ref.ranks)
        dg.close()
    g = ag.Graph.from_edges([], [], n=0)
    dg = device.upload(g)
    run = dg.pagerank(max_iterations=5)
    assert run.iterations == 0 and run.ranks.size == 0
    dg.close()


@pytest.mark.parametrize("fname,fmt", [("graphbrew_graph.bin", ag.FILE_BINARY_EDGE_LIST),
                                       ("graphbrew_graph.org", ag.FILE_TEXT_EDGE_LIST),
                                       ("graphbrew_graph.bin.txt", ag.FILE_TEXT_EDGE_LIST)])
def test_reference_graph_files(device, tmp_path, fname, fmt):
    """The reference's test graph (04_test_graphs/TEST/graphbrew/) read in each of its file
    formats, and again after a serialised-CSR round trip, through the path."""
    say(f"reference graph file {fname}")
    g = ag.Graph.load_file(os.path.join(HERE, "golden", fname), fmt)
    g.save(str(tmp_path / "g.csr"))
    for h in (g, ag.Graph.load_serialized(str(tmp_path / "g.csr"))):
        for pname in ("default", "blocked-all"):
            dg = upload(device, h, pname)
            run = dg.pagerank(kind=32, max_iterations=100, epsilon=1e-6, scaled=True)
            ref = ref_pr(h, kind=32, max_iterations=100, epsilon=1e-6)
            assert run.iterations == ref.iterations
            np.testing.assert_array_equal(run.ranks.view(np.uint32), ref.ranks.view(np.uint32))
            run = dg.pagerank(max_iterations=100)
            ref = ref_pr(h, max_iterations=100)
            assert run.iterations == ref.iterations
            np.testing.assert_allclose(run.ranks, ref.ranks, rtol=0, atol=FP32_TOL)
            dg.close()


# ------------------------------------------------------------------ 16-bit storage

@pytest.mark.parametrize("storage", [ag.STORE_BF16, ag.STORE_F16])
@pytest.mark.parametrize("pname", ["default", "blocked"])
def test_pagerank_16bit_storage_refined(device, rmat16, storage, pname):
    """16-bit contributions (FP32 accumulate) until the L1 delta drops below refine_l1, then
    FP32 contributions to L1 <= 1e-6: the result must be as accurate as the plain FP32 run
    against the FP64 oracle."""
    say(f"pagerank 16-bit storage {storage} {pname}")
    g = rmat16
    dg = upload(device, g, pname)
    kw = dict(max_iterations=300, stop_rule=ag.STOP_L1, l1_tolerance=1e-6,
              init=ag.INIT_UNIFORM)
    plain = dg.pagerank(**kw)
    mixed = dg.pagerank(storage=storage, refine_l1=1e-3, **kw)
    assert mixed.error_total <= 1e-6 and mixed.refine_iteration > 0
    assert mixed.iterations > mixed.refine_iteration
    f64 = f64_ref(g)
    e_plain = np.abs(plain.ranks - f64).max()
    e_mixed = np.abs(mixed.ranks - f64).max()
    l1_mixed = np.abs(mixed.ranks - f64).sum()
    say(f"  plain it={plain.iterations} err={e_plain:.3e}; mixed it={mixed.iterations} "
        f"refine@{mixed.refine_iteration} err={e_mixed:.3e} l1={l1_mixed:.3e}")
    assert e_mixed <= max(2 * e_plain, 1e-9)
    assert np.abs(mixed.ranks - plain.ranks).max() <= FP32_TOL
    # the 16-bit phase really ran: its early deltas differ from the FP32 run's
    assert not np.array_equal(mixed.hist_err[:3], plain.hist_err[:3])
    dg.close()


# ------------------------------------------------------------------ many small graphs

def _small_graphs(count, seed=1):
    rng = np.random.default_rng(seed)
    gs = []
    for i in range(count):
        scale = int(rng.integers(4, 12))
        if i % 3 == 2:
            n = 1 << scale
            gs.append(ag.Graph.erdos_renyi(n, 8 * n, seed=100 + i))
        else:
            gs.append(ag.Graph.rmat(scale, seed=100 + i))
    gs.append(ag.Graph.from_edges([], [], n=1))   # a lone vertex
    gs.append(ag.Graph.from_edges([0, 1, 2], [1, 2, 0], n=3))
    return gs


@pytest.mark.parametrize("rule", ["l1", "active"])
def test_batch_matches_oracle_and_single(device, rule):
    say(f"batch {rule}")
    gs = _small_graphs(40)
    b = ag.Batch(device, gs)
    if rule == "l1":
        kw = dict(max_iterations=300, stop_rule=ag.STOP_L1, l1_tolerance=1e-6,
                  init=ag.INIT_UNIFORM)
        okw = dict(max_iterations=300, stop_rule=1, l1_tolerance=1e-6, init=1)
    else:
        kw = dict(max_iterations=100, stop_rule=ag.STOP_ACTIVE, epsilon=1e-4,
                  init=ag.INIT_REFERENCE)
        okw = dict(max_iterations=100, stop_rule=0, epsilon=1e-4, init=0)
    ranks, its, err, ms = b.pagerank(**kw)
    say(f"  {len(gs)} graphs in {ms:.2f} ms")
    for i, g in enumerate(gs):
        ref = ref_pr(g, **okw)
        assert abs(int(its[i]) - ref.iterations) <= 3, (i, its[i], ref.iterations)
        assert np.abs(ranks[i] - ref.ranks).max() <= FP32_TOL
        if g.n and g.in_degree.max() <= 32:   # every row stream-ordered
            if its[i] == ref.iterations:
                np.testing.assert_array_equal(ranks[i], ref.ranks)
        if rule == "l1" and g.n > 1:
            f64 = f64_ref(g)
            assert np.abs(ranks[i] - f64).max() <= 2 * np.abs(ref.ranks - f64).max() + 1e-9
            dg = device.upload(g)
            single = dg.pagerank(**kw)
            assert np.abs(single.ranks - ranks[i]).max() <= FP32_TOL
            dg.close()
    b.close()


def test_batch_hip_graph_replays(device):
    """The captured launch replayed several times, with eager launches in between, gives the
    eager result every time."""
    say("batch hipGraph replays")
    gs = _small_graphs(24, seed=9)
    b = ag.Batch(device, gs)
    kw = dict(max_iterations=200, stop_rule=ag.STOP_L1, l1_tolerance=1e-6, init=ag.INIT_UNIFORM)
    eager, eits, _, _ = b.pagerank(use_hip_graph=0, **kw)
    for rep in range(4):
        got, gits, _, _ = b.pagerank(use_hip_graph=1, **kw)
        for i in range(len(gs)):
            np.testing.assert_array_equal(got[i], eager[i])
        np.testing.assert_array_equal(gits, eits)
        again, _, _, _ = b.pagerank(use_hip_graph=0, **kw)
        for i in range(len(gs)):
            np.testing.assert_array_equal(again[i], eager[i])
    # the stop rule and the start change between replays of one capture; another iteration
    # budget captures again
    got, gits, _, _ = b.pagerank(use_hip_graph=1, max_iterations=200, stop_rule=ag.STOP_ACTIVE,
                                 epsilon=1e-4, init=ag.INIT_REFERENCE)
    eager, eits, _, _ = b.pagerank(use_hip_graph=0, max_iterations=200,
                                   stop_rule=ag.STOP_ACTIVE, epsilon=1e-4,
                                   init=ag.INIT_REFERENCE)
    np.testing.assert_array_equal(gits, eits)
    for i in range(len(gs)):
        np.testing.assert_array_equal(got[i], eager[i])
    got, gits, _, _ = b.pagerank(use_hip_graph=1, max_iterations=3, stop_rule=ag.STOP_NONE,
                                 init=ag.INIT_UNIFORM)
    assert np.all(gits == 3)
    b.close()


@pytest.mark.parametrize("part_edges,heavy_deg", [(64, 32), (7, 4), (100000, 8)])
def test_batch_part_layouts(device, part_edges, heavy_deg):
    """The parts a graph is cut into (tiny parts: most graphs in many parts, single rows
    bigger than a part alone in theirs; huge parts: one per graph) change only who computes
    which rows: with the same heavy-row threshold every row's value is computed the same way,
    so the ranks equal the default layout's bit for bit whenever the run stops at the same
    iteration (the FP64 L1 delta is summed part by part, so it may differ in its last bits);
    every threshold keeps the oracle's bar."""
    say(f"batch part layout part_edges={part_edges} heavy_deg={heavy_deg}")
    gs = _small_graphs(30, seed=21)
    b = ag.Batch(device, gs)
    kw = dict(max_iterations=300, stop_rule=ag.STOP_L1, l1_tolerance=1e-6, init=ag.INIT_UNIFORM)
    base, bits, berr, _ = b.pagerank(heavy_deg=heavy_deg, **kw)
    got, gits, gerr, _ = b.pagerank(part_edges=part_edges, heavy_deg=heavy_deg, **kw)
    assert np.sum(gits != bits) <= 1
    same = gits == bits
    np.testing.assert_allclose(gerr[same], berr[same], rtol=1e-12, atol=1e-300)
    for i, g in enumerate(gs):
        if same[i]:
            np.testing.assert_array_equal(got[i], base[i])
        ref = ref_pr(g, max_iterations=300, stop_rule=1, l1_tolerance=1e-6, init=1)
        assert abs(int(gits[i]) - ref.iterations) <= 3
        assert np.abs(got[i] - ref.ranks).max() <= FP32_TOL
    b.close()


# ------------------------------------------------------------------ personalised PageRank

@pytest.mark.parametrize("storage", [ag.STORE_F32, ag.STORE_BF16, ag.STORE_F16])
@pytest.mark.parametrize("pname", ["default", "small"])
def test_ppr_64_columns(device, rmat12, storage, pname):
    say(f"ppr 64 columns storage={storage} {pname}")
    g = rmat12
    dg = upload(device, g, pname)
    rng = np.random.default_rng(4)
    seeds = np.concatenate([[ag.PPR_UNIFORM], rng.choice(g.n, 63, replace=False)]).astype(np.uint32)
    ranks, colerr, res = dg.ppr(seeds, max_iterations=300, stop_rule=ag.STOP_L1,
                                l1_tolerance=1e-6, storage=storage, refine_l1=1e-3)
    assert res.max_error <= 1e-6 and np.all(colerr <= 1e-6)
    if storage != ag.STORE_F32:
        assert res.refine_iteration > 0
    # column 0 (uniform restart) is plain PageRank
    plain = dg.pagerank(max_iterations=300, stop_rule=ag.STOP_L1, l1_tolerance=1e-6,
                        init=ag.INIT_UNIFORM)
    assert np.abs(ranks[:, 0] - plain.ranks).max() <= FP32_TOL
    worst = 0.0
    for c in [0, 1, 2, 17, 40, 63]:
        f64 = oracle.ppr_f64(g.out_degree, g.inv_idx, g.in_degree, g.inv_src, int(seeds[c]),
                             threads=8)
        e = np.abs(ranks[:, c] - f64).max()
        worst = max(worst, e / max(f64.max(), 1e-30))
        assert e <= FP32_TOL, (c, e)
        # L1 distance to the FP64 solution within 1e-5 in normalised units (the stop rule
        # leaves at most 0.85e-6 / 0.15 = 5.7e-6)
        assert np.abs(ranks[:, c].astype(np.float64) - f64).sum() <= 1e-5
    # every column: the FP64 residual bounds the L1 distance to the FP64 solution
    resid, mass = oracle.ppr_residual(g.out_degree, g.inv_idx, g.in_degree, g.inv_src, seeds,
                                      ranks, threads=8)
    bound = resid / 0.15
    say(f"  iterations={res.iterations} refine@{res.refine_iteration} "
        f"max relative err vs FP64 {worst:.2e}; residual bound max {bound.max():.2e}")
    assert np.all(bound <= 1e-5)
    dg.close()


def test_ppr_hip_graph_replays(device, rmat12):
    """The captured PPR solve replayed several times, with eager solves in between, gives the
    eager result every time; seeds, column count and storage change between replays without a
    new capture, and a different iteration budget captures again."""
    say("ppr hipGraph replays")
    g = rmat12
    dg = upload(device, g, "small")
    rng = np.random.default_rng(12)
    seeds = np.concatenate([[ag.PPR_UNIFORM], rng.choice(g.n, 63, replace=False)]).astype(np.uint32)
    kw = dict(max_iterations=200, stop_rule=ag.STOP_L1, l1_tolerance=1e-6, refine_l1=1e-3)
    for storage in (ag.STORE_BF16, ag.STORE_F32):
        eager, ecol, eres = dg.ppr(seeds, storage=storage, use_hip_graph=0, **kw)
        for rep in range(3):
            got, gcol, gres = dg.ppr(seeds, storage=storage, use_hip_graph=1, **kw)
            np.testing.assert_array_equal(got, eager)
            np.testing.assert_array_equal(gcol, ecol)
            assert gres.iterations == eres.iterations and gres.executed == eres.executed
            assert gres.refine_iteration == eres.refine_iteration
            again, _, _ = dg.ppr(seeds, storage=storage, use_hip_graph=0, **kw)
            np.testing.assert_array_equal(again, eager)
    few = seeds[5:9].copy()
    eager, _, _ = dg.ppr(few, storage=ag.STORE_F16, use_hip_graph=0, **kw)
    got, _, _ = dg.ppr(few, storage=ag.STORE_F16, use_hip_graph=1, **kw)
    np.testing.assert_array_equal(got, eager)
    got, _, r = dg.ppr(few, storage=ag.STORE_F32, use_hip_graph=1, max_iterations=7,
                       stop_rule=ag.STOP_NONE)
    eager, _, _ = dg.ppr(few, storage=ag.STORE_F32, use_hip_graph=0, max_iterations=7,
                         stop_rule=ag.STOP_NONE)
    assert r.executed == 7
    np.testing.assert_array_equal(got, eager)
    dg.close()


# ------------------------------------------------------------------ autotuner

def _iteration_ms(device, dg, steps=20):
    sess = dg.session(max_iterations=steps + 3, stop_rule=ag.STOP_NONE, init=ag.INIT_UNIFORM)
    sess.step(3)
    device.sync()
    sess.step(steps)
    device.sync()
    ms, _ = sess.last_step_kernel_ms()
    sess.end()
    return ms


def test_autotuner_choice_keeps_results(device, tmp_path):
    """The tuned plan runs the same computation: fixed-point bit-exact and FP32 within
    tolerance of the oracle; the table answers the second call; and timed again here, apart
    from the tuner's own trials, the tuned plan is no slower than the default one (15% for
    noise between runs of ~20 us iterations)."""
    from accelgraph_amd import autotune
    say("autotuner")
    g = ag.Graph.rmat(14)
    table = str(tmp_path / "tuning.json")
    params, rep = autotune.autotune(device, g, iterations=3, repeats=3, passes=1,
                                    table_path=table)
    assert not rep["cached"] and rep["ms_per_iteration"] <= rep["default_ms_per_iteration"]
    params2, rep2 = autotune.autotune(device, g, table_path=table)
    assert rep2["cached"] and autotune._params_dict(params2) == autotune._params_dict(params)
    say(f"  tuned {autotune._params_dict(params)}: {rep['ms_per_iteration']:.4f} ms vs default "
        f"{rep['default_ms_per_iteration']:.4f} ms over {rep['trials']} trials")
    dg = device.upload(g, params=params)
    base = device.upload(g)
    t_tuned = min(_iteration_ms(device, dg) for _ in range(5))
    t_default = min(_iteration_ms(device, base) for _ in range(5))
    base.close()
    say(f"  re-timed: tuned {t_tuned:.4f} ms, default {t_default:.4f} ms per iteration")
    assert t_tuned <= 1.15 * t_default
    fx = dg.pagerank(kind=32, max_iterations=60, epsilon=1e-4)
    ref = ref_pr(g, kind=32, max_iterations=60, epsilon=1e-4)
    np.testing.assert_array_equal(fx.ranks.view(np.uint32), ref.ranks.view(np.uint32))
    f = dg.pagerank(max_iterations=100, epsilon=1e-4)
    rf = ref_pr(g, max_iterations=100, epsilon=1e-4)
    assert np.abs(f.ranks - rf.ranks).max() <= FP32_TOL
    dg.close()


# ------------------------------------------------------------------ learned warm start

def _mlp_case(seed, S):
    from oracle import mlp as omlp
    rng = np.random.default_rng(seed)
    X = np.column_stack([rng.normal(2, 1, S), rng.normal(2, 1, S), rng.normal(3, 1, S),
                         rng.normal(0, 1, S), rng.normal(0, 1, S)]).astype(np.float32)
    y = (0.7 * X[:, 0] - 0.2 * X[:, 3] + 0.5 * X[:, 4]).astype(np.float32)
    p = (rng.normal(size=omlp.P_COUNT) * 0.15).astype(np.float32)
    p[omlp.P_W1 + 5 * omlp.H:omlp.P_B1] = 0.0
    norm = (X.mean(0), X.std(0), float(y.mean()), float(y.std()))
    return X, y, p, norm


@pytest.mark.parametrize("S", [16, 100, 1000])
def test_mlp_gradient_matches_oracle(device, S):
    """Forward loss and every gradient of the BF16-MFMA kernels against the FP64 oracle that
    applies the same BF16 operand roundings."""
    from oracle import mlp as omlp
    say(f"mlp gradient S={S}")
    X, y, p, norm = _mlp_case(S, S)
    m = ag.MLP(device)
    m.set_params(p)
    m.set_normalization(*norm)
    loss, g = m.loss_grad(X, y)
    oloss, og, oyhat = omlp.forward_backward(p, X, y, *norm, bf16_ops=True)
    assert abs(loss - oloss) <= 1e-4 * oloss + 1e-7
    err = np.linalg.norm(g - og) / np.linalg.norm(og)
    say(f"  loss {loss:.6f} vs {oloss:.6f}; relative gradient error {err:.2e}")
    assert err <= 5e-3
    big = np.abs(og) > 1e-3 * np.abs(og).max()
    np.testing.assert_allclose(g[big], og[big], rtol=5e-2)
    pred = m.predict(X)
    np.testing.assert_allclose(pred, oyhat * norm[3] + norm[2], rtol=1e-3, atol=1e-3 * norm[3])
    m.close()


def test_mlp_training_reduces_loss(device):
    say("mlp training")
    X, y, p, norm = _mlp_case(7, 20000)
    m = ag.MLP(device, seed=3)
    m.fit_normalization(X, y)
    l0, _ = m.loss_grad(X, y)
    losses = m.train(X, y, epochs=6, batch=1024, lr=3e-3, seed=1)
    l1, _ = m.loss_grad(X, y)
    say(f"  loss {l0:.4f} -> {l1:.4f} (epochs {losses})")
    assert l1 < 0.1 * l0 and losses[-1] < losses[0]
    m.close()


def test_features_match_definition(device, rmat12):
    """The GPU features against their definition computed from the graph and the oracle's
    first two FP32 iterations."""
    say("warm-start features")
    g = rmat12
    dg = device.upload(g)
    X = dg.features()
    dg.close()
    np.testing.assert_allclose(X[:, 0], np.log1p(g.in_degree.astype(np.float64)), rtol=1e-6)
    np.testing.assert_allclose(X[:, 1], np.log1p(g.out_degree.astype(np.float64)), rtol=1e-6)
    indeg = g.in_degree.astype(np.float64)
    nb = np.zeros(g.n)
    for v in range(g.n):
        b, d = g.inv_idx[v], g.in_degree[v]
        if d:
            nb[v] = np.log1p(indeg[g.inv_src[b:b + d]].mean())
    np.testing.assert_allclose(X[:, 2], nb, rtol=1e-5, atol=1e-6)
    r1 = ref_pr(g, max_iterations=1, stop_rule=2, init=1).scaled
    r2 = ref_pr(g, max_iterations=2, stop_rule=2, init=1).scaled
    np.testing.assert_allclose(X[:, 3], np.log(r1), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(X[:, 4], np.log(r2), rtol=1e-5, atol=1e-6)


def test_warm_start_gain_and_fallback(device):
    """A network trained (on the GPU) on small graphs' solutions starts a held-out graph closer
    to its fixed point: fewer iterations to L1 <= 1e-6 at the same accuracy. An untrained
    network is rejected by the fallback rule and the run still converges."""
    from accelgraph_amd import warmstart
    say("warm start gain")
    m, losses, _ = warmstart.train(device, count=16, lo=9, hi=12, epochs=6, batch=4096,
                                   lr=3e-3, log=say)
    g = ag.Graph.rmat(14, seed=999)
    f64 = f64_ref(g)
    res = warmstart.evaluate(device, g, m, f64)
    say(f"  {res}")
    assert res["warm"]["fallback"] == 0
    assert res["warm"]["iterations"] < res["uniform"]["iterations"]
    assert res["warm"]["max_abs_err_vs_f64"] <= 2 * res["uniform"]["max_abs_err_vs_f64"] + 1e-9
    bad = ag.MLP(device, seed=5)      # untrained: predicts nonsense ranks
    bad.set_normalization(np.zeros(5), np.ones(5), 40.0, 1.0)
    res2 = warmstart.evaluate(device, g, bad, f64)
    assert res2["warm"]["fallback"] == 1
    assert res2["warm"]["max_abs_err_vs_f64"] <= 2 * res2["uniform"]["max_abs_err_vs_f64"] + 1e-9
    m.close()
    bad.close()
