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
