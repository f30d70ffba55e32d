"""The allocation-free SpMV entry point (ag_spmv_pull_noalloc) against ag_spmv_pull, through
the C-ABI: the same kernels on the same inputs, so the products must be equal bit for bit.
Two calls in a row, the second on a smaller graph (the scratch array keeps the first call's
contents); a graph whose vertex count exceeds the fixed scratch is refused with an error."""
import ctypes as C
import sys

import numpy as np
import pytest

import accelgraph_amd as ag
from accelgraph_amd import _abi

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc(dg, x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.full(dg.n, np.nan, np.float32)
    st = ag.lib().ag_spmv_pull_noalloc(dg.handle, x.ctypes.data_as(C.POINTER(C.c_float)),
                                       y.ctypes.data_as(C.POINTER(C.c_float)))
    return st, y


def _weighted(scale, seed):
    g = ag.Graph.rmat(scale, seed=seed)
    rng = np.random.default_rng(seed)
    g.set_weights(((rng.integers(0, 1 << 24, g.m) + 1) * 2.0 ** -24).astype(np.float32))
    x = (rng.integers(0, 1 << 24, g.n) * 2.0 ** -24).astype(np.float32)
    return g, x


def test_noalloc_spmv_matches_allocating_entry_point(device):
    """RMAT-22 (weighted, source-blocked plan, the largest graph the scratch holds) then
    RMAT-14 (unweighted, regular plan): two allocation-free calls in a row, the second on the
    smaller graph, each equal to ag_spmv_pull on the same graph and vector."""
    say("no-alloc SpMV: RMAT-22 weighted, then RMAT-14 unweighted")
    g1, x1 = _weighted(22, 29)
    g2, x2 = _weighted(14, 21)
    d1 = device.upload(g1, weights=True)
    d2 = device.upload(g2, weights=False)
    assert d1.n * 8 <= _abi_scratch_bytes() and d1.info().src_block_bits > 0
    st1, y1 = _noalloc(d1, x1)
    assert st1 == _abi.AG_OK, ag.lib().ag_last_error()
    st2, y2 = _noalloc(d2, x2)
    assert st2 == _abi.AG_OK, ag.lib().ag_last_error()
    want1, _ = d1.spmv(x1)
    want2, _ = d2.spmv(x2)
    np.testing.assert_array_equal(y1.view(np.uint32), want1.view(np.uint32))
    np.testing.assert_array_equal(y2.view(np.uint32), want2.view(np.uint32))
    # and once more on the first graph with other values, after the allocating calls
    x3 = (x1[::-1] * np.float32(0.5)).copy()
    st3, y3 = _noalloc(d1, x3)
    assert st3 == _abi.AG_OK
    np.testing.assert_array_equal(y3.view(np.uint32), d1.spmv(x3)[0].view(np.uint32))
    for h in (d1, d2, g1, g2):
        h.close()


def _abi_scratch_bytes():
    return 32 << 20


def test_noalloc_spmv_refuses_too_many_vertices(device):
    say("no-alloc SpMV: a graph above the scratch limit is refused")
    n = _abi_scratch_bytes() // 8 + 1
    g = ag.Graph.from_edges([0, 1, 2], [1, 2, 0], n=n)
    dg = device.upload(g)
    st, y = _noalloc(dg, np.ones(n, np.float32))
    assert st == _abi.AG_ERR_INVALID
    assert b"scratch" in ag.lib().ag_last_error()
    assert np.isnan(y).all()
    dg.close()
    g.close()
