"""The drop-in entry points themselves (pageRank.h, SPMV.h), called as the reference's callers
call them: a struct Arguments and a struct GraphCSR in, a stats record out. Every dispatch
code of pageRankGraphCSR (pageRank.c:874-939) and SPMVGraphCSR (SPMV.c:594-621) that this
library provides is checked against the oracle: integer kinds bit-exact, FP32 within 1e-5.
The codes this library does not provide (5, 7, 8) print a message and exit the process, so
they are not called here; tests/test_c_caller.py runs them in a child process."""
import ctypes as C
import sys

import numpy as np
import pytest

import accelgraph_amd as ag
from accelgraph_amd import _abi
import oracle

pytestmark = pytest.mark.gpu
FP32_TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def graph():
    g = ag.Graph.rmat(12)
    yield g
    ag.lib().accelgraphShutdown()


def arguments(pushpull, iterations=100, epsilon=1e-4):
    a = _abi.Arguments()
    a.iterations = iterations
    a.epsilon = epsilon
    a.pushpull = pushpull
    a.ker_numThreads = 1
    a.afu_config = a.afu_config_2 = a.cu_config = a.cu_config_2 = 1
    return a


# pushpull code -> (oracle kind, the direct entry point of that variant)
PR_CODES = {0: (0, "pageRankPullGraphCSR"),
            2: (64, "pageRankPullFixedPoint64BitGraphCSR"),
            9: (32, "pageRankPullFixedPoint32BitGraphCSR"),
            10: (16, "pageRankPullFixedPoint16BitGraphCSR"),
            11: (8, "pageRankPullFixedPoint8BitGraphCSR"),
            4: (oracle.KIND_QUANT32, "pageRankPullQuant32BitGraphCSR"),
            12: (oracle.KIND_QUANT16, "pageRankPullQuant16BitGraphCSR"),
            13: (oracle.KIND_QUANT8, "pageRankPullQuant8BitGraphCSR"),
            1: (0, "pageRankPushGraphCSR"),
            3: (64, "pageRankPushFixedPointGraphCSR"),
            6: ("dd", "pageRankDataDrivenPullGraphCSR")}


@pytest.mark.parametrize("pushpull", list(PR_CODES))
def test_pagerank_entry_points(graph, pushpull):
    kind, direct = PR_CODES[pushpull]
    say(f"pageRankGraphCSR pushpull={pushpull} ({direct})")
    L = ag.lib()
    g = graph
    if kind == "dd":
        # summed in the reference's edge order on this unblocked graph: exact
        ref = oracle.pagerank_dd_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                                      g.edges_idx, g.out_degree, g.fwd_dest,
                                      max_iterations=100, epsilon=1e-4)
    else:
        ref = oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src, kind=kind,
                                   max_iterations=100, epsilon=1e-4)
    for fn in (L.pageRankGraphCSR, getattr(L, direct)):
        a = arguments(pushpull)
        st = fn(C.byref(a), g.ptr)
        s = st.contents
        ranks = np.ctypeslib.as_array(s.pageRanks, shape=(g.n,)).copy()
        assert s.num_vertices == g.n
        if kind == 0 and pushpull != 6:
            # the FP32 bar of the parity suite: rows summed in a tree order (pull) or in the
            # LDS atomics' order (push) may move the active-rule stop by a few iterations, ranks
            # agree to 1e-5
            if pushpull == 0:
                assert abs(int(s.iterations) - int(ref.iterations)) <= 3
            else:   # push: stop inside the oracle's tail (ulp-level deltas)
                k = min(int(s.iterations), ref.executed - 1)
                assert ref.hist_active[k] <= max(1, g.n // 1000)
            assert np.abs(ranks - ref.ranks).max() <= FP32_TOL
            np.testing.assert_allclose(ranks, ref.ranks, rtol=5e-5)
            if s.iterations == ref.iterations:   # same final iteration: same L1 delta
                assert abs(s.error_total - ref.error_total) <= 1e-3 * ref.error_total + 1e-7
        else:
            assert s.iterations == ref.iterations and s.activeVertices == ref.active
            np.testing.assert_array_equal(ranks.view(np.uint32), ref.ranks.view(np.uint32))
            assert abs(s.error_total - ref.error_total) <= 1e-12 + 1e-9 * ref.error_total
        assert s.damp == np.float32(0.85) and s.base_pr == np.float32(1) - np.float32(0.85)
        L.freePageRankStats(st)


def test_spmv_entry_points(graph):
    say("SPMVGraphCSR pushpull 0, 2, 1 and 3")
    L = ag.lib()
    g = graph
    x = np.where(g.out_degree > 0, np.float32(1) / np.maximum(g.out_degree, 1).astype(np.float32),
                 np.float32(0.001)).astype(np.float32)
    y_ref = oracle.spmv_f32(g.inv_idx, g.in_degree, g.inv_src, None, x)
    a = arguments(0, iterations=3)
    st = L.SPMVGraphCSR(C.byref(a), g.ptr)
    s = st.contents
    assert s.iterations == 3
    np.testing.assert_array_equal(np.ctypeslib.as_array(s.vector_input, shape=(g.n,)), x)
    y = np.ctypeslib.as_array(s.vector_output, shape=(g.n,)).copy()
    np.testing.assert_allclose(y, y_ref, rtol=1e-6, atol=0)
    L.freeSPMVStats(st)

    xq = oracle.fixed_encode(32, x).astype(np.uint32)
    wq = int(oracle.fixed_encode(32, np.float32([0.0001]))[0])
    yq = oracle.spmv_fixed(g.inv_idx, g.in_degree, g.inv_src, None, wq, xq)
    a = arguments(2, iterations=2)
    st = L.SPMVGraphCSR(C.byref(a), g.ptr)
    y = np.ctypeslib.as_array(st.contents.vector_output, shape=(g.n,)).copy()
    np.testing.assert_array_equal(y, oracle.fixed_decode(32, yq).astype(np.float32))
    L.freeSPMVStats(st)

    # push (1): three accumulated products; fixed push (3): Q32.32, two products
    fwd = (g.edges_idx, g.out_degree, g.fwd_dest)
    a = arguments(1, iterations=3)
    st = L.SPMVGraphCSR(C.byref(a), g.ptr)
    y = np.ctypeslib.as_array(st.contents.vector_output, shape=(g.n,)).copy()
    ref = oracle.spmv_push_f32(*fwd, None, x, reps=3)
    bound = (3 * g.in_degree.astype(np.float64) + 2) * 2.0 ** -24 * 2 * np.abs(ref)
    assert st.contents.iterations == 3
    assert np.all(np.abs(y.astype(np.float64) - ref) <= bound + 1e-30)
    L.freeSPMVStats(st)
    a = arguments(3, iterations=2)
    st = L.SPMVGraphCSR(C.byref(a), g.ptr)
    y = np.ctypeslib.as_array(st.contents.vector_output, shape=(g.n,)).copy()
    yq = oracle.spmv_push_fixed(*fwd, None, oracle.fixed_encode(64, x), reps=2)
    np.testing.assert_array_equal(y, oracle.fixed_decode(64, yq).astype(np.float32))
    L.freeSPMVStats(st)


def test_graph_rebuilt_in_the_same_buffers():
    """The reference's vertex-order sweep frees a graph and builds the same graph relabelled,
    with the same N and M, in a loop, and never releases anything
    (02_capi_graph/src/tests/sweep_order-PR-performance-graph.c:205-233); freed arrays of
    that size come back at the same addresses. Here the relabelled graph is written into the
    very same buffers: the drop-in call must notice the new contents and return the new
    graph's ranks (fixed-32 bit-exact with the oracle), twice in a row, then again after the
    first graph is written back."""
    say("drop-in cache: relabelled graph written into the same buffers")
    L = ag.lib()
    g = ag.Graph.rmat(12, seed=4242)
    rng = np.random.default_rng(17)
    perm = rng.permutation(g.n).astype(np.uint32)
    fsrc = np.repeat(np.arange(g.n, dtype=np.uint32), g.out_degree.astype(np.int64))
    h = ag.Graph.from_edges(perm[fsrc], perm[g.fwd_dest], n=g.n)
    assert h.n == g.n and h.m == g.m
    views = lambda x: (x.out_degree, x.edges_idx, x.fwd_src, x.fwd_dest, x.in_degree,
                       x.inv_idx, x.inv_src)
    first = [a.copy() for a in views(g)]

    def run_fixed32():
        st = L.pageRankPullFixedPoint32BitGraphCSR(C.byref(arguments(9)), g.ptr)
        ranks = np.ctypeslib.as_array(st.contents.pageRanks, shape=(g.n,)).copy()
        its = int(st.contents.iterations)
        L.freePageRankStats(st)
        return ranks, its

    def expect(x):
        return oracle.pagerank_pull(x.out_degree, x.inv_idx, x.in_degree, x.inv_src, kind=32,
                                    max_iterations=100, epsilon=1e-4)

    want_g = expect(g)
    want_h = expect(h)
    assert not np.array_equal(want_g.ranks, want_h.ranks)
    for contents, want in ((first, want_g), (views(h), want_h), (views(h), want_h),
                           (first, want_g)):
        for dst, src in zip(views(g), contents):
            dst[:] = src
        ranks, its = run_fixed32()
        assert its == want.iterations
        np.testing.assert_array_equal(ranks.view(np.uint32), want.ranks.view(np.uint32))
    L.accelgraphShutdown()
    g.close()
    h.close()


def test_reference_order_through_the_drop_in(graph, monkeypatch):
    """ACCELGRAPH_FP32_ORDER=reference: pageRankPullGraphCSR sums every row in the reference's
    order, so its iteration count and ranks equal the sequential oracle's bit for bit (the
    default fast order may stop a few iterations apart)."""
    say("pageRankPullGraphCSR with ACCELGRAPH_FP32_ORDER=reference")
    L = ag.lib()
    g = graph
    monkeypatch.setenv("ACCELGRAPH_FP32_ORDER", "reference")
    ref = oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src,
                               max_iterations=100, epsilon=1e-4)
    st = L.pageRankPullGraphCSR(C.byref(arguments(0)), g.ptr)
    s = st.contents
    ranks = np.ctypeslib.as_array(s.pageRanks, shape=(g.n,)).copy()
    assert s.iterations == ref.iterations and s.activeVertices == ref.active
    np.testing.assert_array_equal(ranks.view(np.uint32), ref.ranks.view(np.uint32))
    assert abs(s.error_total - ref.error_total) <= 1e-12 + 1e-9 * ref.error_total
    L.freePageRankStats(st)
    monkeypatch.delenv("ACCELGRAPH_FP32_ORDER")
    L.accelgraphReleaseGraph()
