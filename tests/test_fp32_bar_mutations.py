"""The per-row FP32 bars (tests/fp32_bars.py) have teeth: mistakes planted in CPU evaluations
of the gather fail them, while every correct evaluation order passes.

The planted mistakes follow the reference's own mutation suite for the PageRank sum kernel
(03_capi_integration/accelerator_verification/rtl/unit/algorithms/suites.json:64-170:
no-accumulate, wrong write index, edge counter stepping by two, a truncated term) plus the
failure modes of this library's schedule: one term dropped from every row, one term dropped
from a single row, and one 16384-edge segment (a hub-row segment or a source-block piece)
dropped from a heavy row. All of this runs on the CPU; nothing here touches a GPU.
"""
import numpy as np
import pytest

import accelgraph_amd as ag
import oracle
from fp32_bars import (BASE, DAMP, anchor_check, anchor_violations, contributions,
                       one_step_bound, one_step_check)


def _update(sums):
    return (BASE + DAMP * np.asarray(sums, np.float32)).astype(np.float32)


def _rows(g):
    deg = g.in_degree.astype(np.int64)
    return np.repeat(np.arange(g.n), deg), deg


def _mutant_sums(g, c, mutation):
    """FP32 row sums of contributions c with a planted mistake (sums are formed in FP64 and
    rounded once, so only the mistake separates them from a correct evaluation)."""
    row, deg = _rows(g)
    vals = c.astype(np.float64)[g.inv_src]
    start = np.repeat(np.cumsum(deg) - deg, deg)
    k = np.arange(vals.size) - start          # position of each edge inside its row
    keep = np.ones(vals.size, bool)
    if mutation == "drop-last-term":
        keep &= ~((k == np.repeat(deg, deg) - 1) & (np.repeat(deg, deg) >= 2))
    elif mutation == "no-accumulate":          # only the last term survives
        keep &= k == np.repeat(deg, deg) - 1
    elif mutation == "edge-counter-by-two":    # every other edge skipped
        keep &= (k % 2) == 0
    elif mutation == "truncated-term":         # the last term halved
        last = (k == np.repeat(deg, deg) - 1) & (np.repeat(deg, deg) >= 2)
        vals = np.where(last, vals * 0.5, vals)
    elif mutation == "drop-one-term-heaviest-row":
        h = int(np.argmax(deg))
        keep &= ~((row == h) & (k == deg[h] // 2))
    elif mutation == "drop-segment-heavy-rows":
        heavy = np.repeat(deg, deg) > 16384
        keep &= ~(heavy & (k >= 16384) & (k < 2 * 16384))
    s = np.bincount(row[keep], weights=vals[keep], minlength=g.n)
    if mutation == "wrong-write-index":        # row v's sum lands in row v + 1
        s = np.roll(s, 1)
    return s.astype(np.float32)


@pytest.fixture(scope="module")
def graphs():
    rng = np.random.default_rng(3)
    g14 = ag.Graph.rmat(14)
    # a graph with a 40000-in-edge row (several 16384-edge segments) and an RMAT-like rest
    n = 50000
    src = np.concatenate([np.arange(1, 40001), rng.integers(0, n, 200000)])
    dst = np.concatenate([np.zeros(40000, np.int64), rng.integers(0, n, 200000)])
    hub = ag.Graph.from_edges(src.astype(np.uint32), dst.astype(np.uint32), n=n)
    return {"rmat14": g14, "hub": hub}


def _init(g, seed=1):
    return (np.random.default_rng(seed).random(g.n) * 2 + 0.05).astype(np.float32)


@pytest.mark.parametrize("gname", ["rmat14", "hub"])
def test_correct_orders_pass(graphs, gname):
    """The oracle's sequential sum, the same sum over rows reversed, and the correctly rounded
    sum all pass the one-iteration bar."""
    g = graphs[gname]
    init = _init(g)
    seq = oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                               max_iterations=1, stop_rule=2, init=2, init_ranks=init)
    one_step_check(g, init, seq.scaled, "oracle")
    rsrc = g.inv_src.copy()
    for v in np.nonzero(g.in_degree > 1)[0]:
        b, d = g.inv_idx[v], g.in_degree[v]
        rsrc[b:b + d] = rsrc[b:b + d][::-1]
    rev = oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, rsrc,
                               max_iterations=1, stop_rule=2, init=2, init_ranks=init)
    one_step_check(g, init, rev.scaled, "oracle reversed")
    c = contributions(g.out_degree, init)
    one_step_check(g, init, _update(_mutant_sums(g, c, "none")), "correctly rounded")


MUTATIONS = ["drop-last-term", "no-accumulate", "edge-counter-by-two", "truncated-term",
             "wrong-write-index", "drop-one-term-heaviest-row"]


@pytest.mark.parametrize("mutation", MUTATIONS)
def test_one_step_bar_rejects_mutants(graphs, mutation):
    g = graphs["rmat14"]
    init = _init(g)
    c = contributions(g.out_degree, init)
    got = _update(_mutant_sums(g, c, mutation))
    with pytest.raises(AssertionError):
        one_step_check(g, init, got, mutation)


def test_one_step_bar_rejects_dropped_hub_segment(graphs):
    g = graphs["hub"]
    init = _init(g)
    c = contributions(g.out_degree, init)
    got = _update(_mutant_sums(g, c, "drop-segment-heavy-rows"))
    expected, bound = one_step_bound(g.out_degree, g.inv_idx, g.in_degree, g.inv_src, init)
    # only the hub row is wrong, and it alone fails
    bad = np.nonzero(np.abs(got.astype(np.float64) - expected) > bound)[0]
    assert bad.tolist() == [0]


def _iterate(g, iterations, mutation="none"):
    r = np.ones(g.n, np.float32)
    for _ in range(iterations):
        r = _update(_mutant_sums(g, contributions(g.out_degree, r), mutation))
    return r / np.float32(g.n)


@pytest.mark.parametrize("mutation", ["none", "drop-one-term-heaviest-row", "drop-last-term",
                                      "truncated-term"])
def test_anchor_bar(graphs, mutation):
    """After 20 iterations from the uniform start: the correctly rounded evaluation passes the
    FP64 anchor; each mutant fails it (a single dropped term in the heaviest row included)."""
    g = graphs["rmat14"]
    ora = oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                               max_iterations=20, stop_rule=2, init=1)
    f64 = oracle.pagerank_f64(g.out_degree, g.inv_idx, g.in_degree, g.inv_src, tol=0.0,
                              max_iterations=20)
    got = _iterate(g, 20, mutation)
    if mutation == "none":
        anchor_check(got, ora.ranks, f64)
    else:
        assert anchor_violations(got, ora.ranks, f64).size > 0
        # the former bar (1e-5 max-abs on normalised ranks) let this mutant through
        if mutation == "drop-one-term-heaviest-row":
            assert np.abs(got - ora.ranks).max() <= 1e-5
