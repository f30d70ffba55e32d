"""Invariants of the degree-binned pull plan (host side, plan.c). The GPU kernels index
tiles, segments and rows only through these arrays, so checking them on the CPU bounds
every device access."""
import ctypes as C

import numpy as np
import pytest

import accelgraph_amd as ag
from accelgraph_amd._abi import PlanParams

u32p = C.POINTER(C.c_uint32)


class Plan(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint32), ("params", PlanParams),
                ("d2o", u32p), ("o2d", u32p), ("off", u32p), ("src", u32p),
                ("w", C.POINTER(C.c_float)), ("outdeg", u32p), ("bin_row", C.c_uint32 * 6),
                ("bin_edges", C.c_uint64 * 5), ("n_tiles", C.c_uint32), ("tile_row", u32p),
                ("n_hubs", C.c_uint32), ("n_segs", C.c_uint32), ("seg_row", u32p),
                ("seg_lo", u32p), ("seg_hi", u32p), ("hub_seg0", u32p),
                ("bin_blocks", C.c_uint32 * 5), ("slot_blk0", C.c_uint32 * 6),
                ("total_blocks", C.c_uint32)]


def build(g, params=None, n=None, outdeg=None, idx=None, deg=None, src=None):
    L = ag.lib()
    L.ag_plan_build.restype = C.c_int
    L.ag_plan_build.argtypes = [C.c_uint32, C.c_uint32, u32p, u32p, u32p, u32p,
                                C.POINTER(C.c_float), C.POINTER(PlanParams), C.POINTER(Plan),
                                C.POINTER(C.c_char_p)]
    L.ag_plan_free.argtypes = [C.POINTER(Plan)]
    if g is not None:
        n, outdeg, idx, deg, src = g.n, g.out_degree, g.inv_idx, g.in_degree, g.inv_src
    arrs = [np.ascontiguousarray(a, np.uint32) for a in (outdeg, idx, deg, src)]
    p = Plan()
    err = C.c_char_p()
    st = L.ag_plan_build(n, arrs[3].size, *[a.ctypes.data_as(u32p) for a in arrs], None,
                         C.byref(params) if params else None, C.byref(p), C.byref(err))
    return st, p, err.value


def view(ptr, k):
    return np.ctypeslib.as_array(ptr, shape=(k,)).copy() if k else np.zeros(0, np.uint32)


def check_plan(g, p):
    n, prm = g.n, p.params
    d2o, o2d = view(p.d2o, n), view(p.o2d, n)
    assert sorted(d2o) == list(range(n))
    np.testing.assert_array_equal(o2d[d2o], np.arange(n))
    off = view(p.off, n + 1)
    srcd = view(p.src, p.m)
    deg = np.diff(off)
    np.testing.assert_array_equal(deg, g.in_degree[d2o])
    np.testing.assert_array_equal(view(p.outdeg, n), g.out_degree[d2o])
    # every device row reads the same sources, in the same order, relabelled
    for r in range(n):
        v = d2o[r]
        orig = g.inv_src[g.inv_idx[v]:g.inv_idx[v] + g.in_degree[v]]
        np.testing.assert_array_equal(d2o[srcd[off[r]:off[r + 1]]], orig)
    # bins are contiguous and respect thresholds
    br = list(p.bin_row)
    assert br[0] == 0 and br[-1] == n and br == sorted(br)
    lims = [prm.thread_max_deg, prm.group16_max_deg, prm.wave_max_deg, prm.block_max_deg,
            2 ** 32]
    lows = [-1] + lims[:-1]
    for b in range(5):
        d = deg[br[b]:br[b + 1]]
        assert np.all(d <= lims[b]) and np.all(d > lows[b])
    # tiles: cover the thread bin exactly, bounded rows and in-edges
    tr = view(p.tile_row, p.n_tiles + 1)
    assert tr[0] == br[0] and tr[-1] == br[1] and np.all(np.diff(tr) >= 1)
    assert np.all(np.diff(tr) <= prm.tile_rows)
    assert np.all(off[tr[1:]] - off[tr[:-1]] <= prm.tile_nnz)
    assert prm.tile_nnz <= 4096 and prm.tile_rows <= 1024
    # hub segments: consecutive, cover every hub row, bounded length
    rows, lo, hi = view(p.seg_row, p.n_segs), view(p.seg_lo, p.n_segs), view(p.seg_hi, p.n_segs)
    s0 = view(p.hub_seg0, p.n_hubs + 1)
    assert p.n_hubs == br[5] - br[4] and s0[-1] == p.n_segs
    for h in range(p.n_hubs):
        r = br[4] + h
        segs = range(s0[h], s0[h + 1])
        assert len(segs) >= 1
        assert lo[segs[0]] == off[r] and hi[segs[-1]] == off[r + 1]
        for s in segs:
            assert rows[s] == r and 0 < hi[s] - lo[s] <= prm.segment_nnz
        for a, b in zip(segs[:-1], segs[1:]):
            assert hi[a] == lo[b]
    # launch slots: HUB, BLOCK, WAVE, GROUP16, THREAD
    blk = list(p.bin_blocks)
    assert blk == [p.n_tiles, -(-(br[2] - br[1]) // 16), -(-(br[3] - br[2]) // 4),
                   br[4] - br[3], p.n_segs]
    sl = list(p.slot_blk0)
    assert sl == [0, blk[4], blk[4] + blk[3], blk[4] + blk[3] + blk[2],
                  blk[4] + blk[3] + blk[2] + blk[1], sum(blk)]
    assert p.total_blocks == sum(blk)


SMALL = dict(thread_max_deg=2, group16_max_deg=6, wave_max_deg=20, block_max_deg=60,
             tile_rows=7, tile_nnz=9, segment_nnz=256, order=0, cold_below_bin=0,
             wg_per_cu=0)


@pytest.mark.parametrize("params", [None, SMALL, dict(SMALL, order=1)],
                         ids=["default", "small", "small-outdeg-order"])
def test_plan_invariants_rmat(params):
    g = ag.Graph.rmat(11)
    prm = ag.plan_params(**params) if params else None
    st, p, err = build(g, prm)
    assert st == 0, err
    try:
        check_plan(g, p)
        if params:
            assert all(p.bin_row[b + 1] > p.bin_row[b] for b in range(5)), "every bin used"
    finally:
        ag.lib().ag_plan_free(C.byref(p))


def test_plan_empty_graph():
    g = ag.Graph.from_edges([], [], n=0)
    st, p, err = build(g)
    assert st == 0 and p.total_blocks == 0
    ag.lib().ag_plan_free(C.byref(p))


def test_plan_rejects_bad_input():
    g = ag.Graph.rmat(6)
    bad = g.inv_src.copy()
    bad[0] = g.n  # source id out of range
    st, p, err = build(None, n=g.n, outdeg=g.out_degree, idx=g.inv_idx, deg=g.in_degree,
                       src=bad)
    assert st == 1 and b"source" in err
    idx = g.inv_idx.copy()
    idx[-1] = g.m  # row runs past num_edges
    deg = g.in_degree.copy()
    deg[-1] = max(1, deg[-1])
    st, p, err = build(None, n=g.n, outdeg=g.out_degree, idx=idx, deg=deg, src=g.inv_src)
    assert st == 1
    st, p, err = build(g, ag.plan_params(tile_nnz=5000))
    assert st == 1
    st, p, err = build(g, ag.plan_params(thread_max_deg=100, group16_max_deg=50))
    assert st == 1
