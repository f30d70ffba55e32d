"""Invariants of the degree-binned pull plan (host side, plan.c). The GPU kernels index
tiles, segments and rows only through these arrays, so checking them on the CPU bounds
every device access."""
import ctypes as C

import numpy as np
import pytest

import accelgraph_amd as ag
from accelgraph_amd._abi import PlanParams

u32p = C.POINTER(C.c_uint32)


class BlockPlan(C.Structure):
    _fields_ = [("bits", C.c_uint32), ("nblocks", C.c_uint32), ("row0", C.c_uint32),
                ("nv", C.c_uint32), ("mb", C.c_uint32), ("slice_row", C.c_uint32 * 9),
                ("voff", u32p), ("vsrc", u32p), ("vw", C.POINTER(C.c_float)),
                ("n_items", C.c_uint32), ("item", u32p), ("item_off", C.c_uint32 * 9),
                ("coff", u32p), ("cidx", u32p), ("lpos", C.POINTER(C.c_uint16)),
                ("c_n_tiles", C.c_uint32), ("ct_row", u32p), ("ct_start", u32p),
                ("ct_pre", u32p), ("c_n_items", C.c_uint32), ("c_item", u32p),
                ("c_item_off", C.c_uint32 * 9), ("r_slot_blk0", C.c_uint32 * 6),
                ("hot_bits", C.c_uint32), ("n_hot", C.c_uint32), ("hot_edges", C.c_uint64),
                ("hot_seg", u32p), ("hot_end", C.c_uint32 * 8), ("n_hgrp", C.c_uint32),
                ("hgrp", u32p), ("hgrp_off", C.c_uint32 * 9), ("piece", C.c_uint32),
                ("top0", C.c_uint32), ("top_edges", C.c_uint64), ("sorted", C.c_uint32),
                ("vsp", u32p), ("tbase", u32p)]


class Plan(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint32), ("params", PlanParams),
                ("d2o", u32p), ("o2d", u32p), ("off", u32p), ("src", u32p),
                ("w", C.POINTER(C.c_float)), ("outdeg", u32p), ("bin_row", C.c_uint32 * 6),
                ("bin_edges", C.c_uint64 * 5), ("n_tiles", C.c_uint32), ("tile_row", u32p),
                ("n_hubs", C.c_uint32), ("n_segs", C.c_uint32), ("seg_row", u32p),
                ("seg_lo", u32p), ("seg_hi", u32p), ("hub_seg0", u32p),
                ("bin_blocks", C.c_uint32 * 5), ("slot_blk0", C.c_uint32 * 6),
                ("total_blocks", C.c_uint32), ("blk", BlockPlan)]


def build(g, params=None, n=None, outdeg=None, idx=None, deg=None, src=None, weights=None):
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
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    st = L.ag_plan_build(n, arrs[3].size, *[a.ctypes.data_as(u32p) for a in arrs],
                         None if w is None else w.ctypes.data_as(C.POINTER(C.c_float)),
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
            assert rows[s] == r and 0 < hi[s] - lo[s]
            assert hi[s] - lo[s] <= prm.segment_nnz or prm.fp32_order == 1
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
             tile_rows=7, tile_nnz=9, segment_nnz=256, order=0, src_block_bits=255,
             wg_per_cu=0, blocked_min_bin=1)


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


@pytest.mark.parametrize("params", [{}, SMALL, dict(src_block_bits=8, tile_nnz=64)],
                         ids=["default", "small", "blocked-request"])
def test_plan_reference_order(params):
    """fp32_order = reference: rows up to tile_nnz in-edges are thread-bin rows (one lane
    each), every longer row is a hub row of exactly one segment, the middle bins are empty and
    source blocking is off, whatever the other parameters ask."""
    g = ag.Graph.rmat(12)
    st, p, err = build(g, ag.plan_params(**dict(params, fp32_order=1)))
    assert st == 0, err
    try:
        prm = p.params
        assert prm.fp32_order == 1 and prm.thread_max_deg == prm.tile_nnz
        check_plan(g, p)
        assert all(p.bin_row[b + 1] == p.bin_row[b] for b in (1, 2, 3))
        assert p.n_segs == p.n_hubs and p.blk.bits == 0
        deg = np.diff(view(p.off, g.n + 1).astype(np.int64))
        assert np.all(deg[p.bin_row[4]:] > prm.tile_nnz)
        assert np.all(deg[:p.bin_row[1]] <= prm.tile_nnz)
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
    # a vertex count whose index arithmetic would leave 32 bits is refused before any array
    # is read (the arrays here are far shorter than the count)
    st, p, err = build(None, n=0xFFFF0001, outdeg=g.out_degree, idx=g.inv_idx, deg=g.in_degree,
                       src=g.inv_src)
    assert st == 1 and b"num_vertices" in err


# ---------------------------------------------------------------- source blocking

PIECE = 4096
NXCD = 8
ITEM_REG, ITEM_VTILE, ITEM_CTILE, ITEM_CBIG = 0, 1, 2, 3


HOT_NNZ = 2048


def block_keys(B, n):
    """Source block of every device id: hot segment b for the ids of B.hot_seg[b], else the
    ordinary 2^bits block after the hot ones."""
    ids = np.arange(n, dtype=np.int64)
    keys = B.n_hot + (ids >> B.bits)
    if B.n_hot:
        hs = view(B.hot_seg, B.n_hot).astype(np.int64)
        seg = ids >> B.hot_bits
        pos = np.searchsorted(hs, seg)
        hit = (pos < hs.size) & (hs[np.minimum(pos, hs.size - 1)] == seg)
        keys[hit] = pos[hit]
    return keys


def check_sorted_tiles(p, it):
    """tile_sort: with blocks of at most 2^20 sources (and the sort not switched off) every
    pass-1 tile's packed words decode to a permutation of its positions whose sources are
    the tile's ids in stream order, in ascending source order with equal sources in position
    order; tbase is the block start of the tile's sources."""
    B, prm = p.blk, p.params
    want = B.bits <= 20 and prm.tile_sort != ag.SORT_OFF and B.n_hot == 0
    assert B.sorted == int(want)
    if not B.sorted:
        return
    voff = view(B.voff, B.nv + 1).astype(np.int64)
    vsrc = view(B.vsrc, B.mb).astype(np.int64)
    vsp = view(B.vsp, B.mb).astype(np.int64)
    tb = view(B.tbase, B.n_items).astype(np.int64)
    mask = (1 << B.bits) - 1
    for i, (first, kc) in enumerate(it):
        if kc & 15 != ITEM_VTILE:
            continue
        e0, e1 = voff[first], voff[first + (kc >> 4)]
        w = vsp[e0:e1]
        pos, s = w >> B.bits, tb[i] + (w & mask)
        assert tb[i] == (vsrc[e0] >> B.bits) << B.bits
        assert np.array_equal(np.sort(pos), np.arange(e1 - e0))
        assert np.array_equal(s, vsrc[e0 + pos])
        key = s * 4096 + pos
        assert np.all(np.diff(key) > 0)


def check_block_plan(g, p):
    """Every blocked row's in-edges appear exactly once among its virtual rows, grouped by
    source block in stream order; virtual rows are ordered (slice, block, row, piece); the
    pass-1 tiles of each slice cover that slice's virtual rows exactly once, block-major,
    within the tile limits; the combine tiles and big rows cover every blocked row once, each
    tile's per-block runs hold exactly its rows' partials and lpos places them row-major."""
    B, prm, n = p.blk, p.params, g.n
    bits, nb, row0, nv = B.bits, B.nblocks, B.row0, B.nv
    piece = B.piece
    if prm.piece_nnz:
        assert piece == prm.piece_nnz
    else:   # the default limit, or single in-edges when pieces are nearly as many as edges
        assert piece in (min(PIECE, prm.tile_nnz), 1)
        if piece == 1 and min(PIECE, prm.tile_nnz) > 1:
            assert B.n_hot or B.nv == B.mb
    hot_piece = min(HOT_NNZ, prm.tile_nnz)
    # first-pass LDS table: the last 4096 device ids, kept when asked or when they receive at
    # least 2% of the blocked in-edges
    t0 = max(n - 4096, 0)
    te = int(np.count_nonzero(view(p.src, p.m)[view(p.off, n + 1)[B.row0]:] >= t0))
    on = prm.top_table == ag.TOP_ON or (prm.top_table == ag.TOP_AUTO and te * 50 >= B.mb)
    assert (B.top0, B.top_edges) == ((t0, te) if on else (n, 0))
    off = view(p.off, n + 1)
    srcd = view(p.src, p.m)
    br = list(p.bin_row)
    assert row0 == br[prm.blocked_min_bin] and B.mb == off[n] - off[row0]
    H = B.n_hot
    assert nb == H + -(-n // (1 << bits)) and nb <= 256
    blk_of = block_keys(B, n)   # source block of every device id
    hot_seg = view(B.hot_seg, H)
    if H:
        assert B.hot_bits == min(14, bits) and np.all(np.diff(hot_seg.astype(np.int64)) > 0)
        # the hot segments are the qualifying ones with the most in-edges
        kb = B.hot_bits
        rows = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))[off[row0]:]
        seg = srcd[off[row0]:].astype(np.int64) >> kb
        E = np.bincount(seg, minlength=-(-n >> kb) or 1)
        P = np.bincount(np.unique(rows * (1 << 20) + seg) % (1 << 20), minlength=E.size)
        ok = (E >= (4 << kb)) & (2 * P <= E)
        assert set(hot_seg.tolist()) <= set(np.nonzero(ok)[0].tolist())
        assert B.hot_edges == E[hot_seg].sum()
        if H < min(prm.hot_segments, 256 - (nb - H)):
            assert ok.sum() == H
        else:
            assert E[hot_seg].min() >= E[ok & ~np.isin(np.arange(E.size), hot_seg)].max(initial=0)
    else:
        assert B.hot_bits == 0 and B.hot_edges == 0 and B.n_hgrp == 0
    sl = list(B.slice_row)
    assert sl[0] == row0 and sl[-1] == n and sl == sorted(sl)
    voff, vsrc = view(B.voff, nv + 1), view(B.vsrc, B.mb)
    coff, cidx = view(B.coff, n + 1), view(B.cidx, nv)
    lpos = np.ctypeslib.as_array(B.lpos, shape=(nv,)).copy()
    assert voff[0] == 0 and voff[-1] == B.mb
    vlen = np.diff(voff.astype(np.int64))
    assert np.all(vlen >= 1) and np.all(vlen <= piece)
    assert np.all(coff[:row0 + 1] == 0) and coff[n] == nv
    assert np.all(np.diff(coff.astype(np.int64)) >= 0)
    assert sorted(cidx) == list(range(nv))
    vrow = np.empty(nv, np.int64)
    vblk = np.empty(nv, np.int64)
    for r in range(row0, n):
        vs = cidx[coff[r]:coff[r + 1]]
        vrow[vs] = r
        seq = [vsrc[voff[v]:voff[v + 1]] for v in vs]
        for v, sq in zip(vs, seq):
            blks = np.unique(blk_of[sq])
            assert blks.size == 1
            vblk[v] = blks[0]
        got = np.concatenate(seq) if seq else np.zeros(0, np.uint32)
        row = srcd[off[r]:off[r + 1]]
        want = row[np.argsort(blk_of[row], kind="stable")]
        np.testing.assert_array_equal(got, want)
        # pieces: consecutive virtual rows of one block are full pieces except the last
        bl = vblk[vs]
        assert np.all(np.diff(bl) >= 0)
        for k in range(len(vs) - 1):
            if bl[k] == bl[k + 1]:
                assert vlen[vs[k]] == (hot_piece if bl[k] < H else piece)
                assert vs[k + 1] == vs[k] + 1
        assert all(vlen[v] <= hot_piece for v in vs if vblk[v] < H)
    xs = np.searchsorted(np.array(sl[1:]), vrow, side="right")
    key = xs * nb + vblk
    assert np.all(np.diff(key) >= 0), "virtual rows ordered by (slice, block)"
    same = np.diff(key) == 0
    assert np.all(np.diff(vrow)[same] >= 0), "then by row"
    # pass-1 items
    it = view(B.item, 2 * B.n_items).reshape(-1, 2).astype(np.int64)
    io = list(B.item_off)
    assert io[0] == 0 and io[-1] == B.n_items and io == sorted(io)
    covered = np.zeros(nv, np.int64)
    reg_total = B.r_slot_blk0[5]
    reg_seen = []
    hg = view(B.hgrp, 3 * B.n_hgrp).reshape(-1, 3).astype(np.int64)
    ho = list(B.hgrp_off)
    assert ho[0] == 0 and ho[-1] == B.n_hgrp and ho == sorted(ho)
    for x in range(NXCD):
        last_blk = -1
        he = B.hot_end[x]
        assert io[x] <= he <= io[x + 1]
        for i, (first, kc) in enumerate(it[io[x]:io[x + 1]], io[x]):
            kind, cnt = kc & 15, kc >> 4
            if kind == ITEM_REG:
                assert cnt == 0 and first % NXCD == x and i >= he
                reg_seen.append(first)
                continue
            assert kind == ITEM_VTILE
            vs = np.arange(first, first + cnt)
            assert cnt >= 1 and vs[-1] < nv
            covered[vs] += 1
            assert np.all(xs[vs] == x) and np.unique(vblk[vs]).size == 1
            assert vblk[vs[0]] >= last_blk
            last_blk = vblk[vs[0]]
            assert (vblk[vs[0]] < H) == (i < he)
            assert vlen[vs].sum() <= (hot_piece if vblk[vs[0]] < H else prm.tile_nnz)
        # the groups of slice x cover its hot items [io[x], he) once, in order, one segment
        # each, at most 65536 edges
        nxt = io[x]
        for first, cnt, sg in hg[ho[x]:ho[x + 1]]:
            assert first == nxt and cnt >= 1
            nxt += cnt
            vs = np.concatenate([np.arange(it[i, 0], it[i, 0] + (it[i, 1] >> 4))
                                 for i in range(first, first + cnt)])
            assert np.all(vblk[vs] == np.searchsorted(hot_seg, sg)) and hot_seg[vblk[vs[0]]] == sg
            assert vlen[vs].sum() <= 65536
        assert nxt == he
    assert np.all(covered == 1)
    assert sorted(reg_seen) == list(range(reg_total))
    check_sorted_tiles(p, it)
    # regular schedule of the unblocked rows: bins below blocked_min_bin only
    blk_counts = list(p.bin_blocks)
    rb = [blk_counts[b] if b < prm.blocked_min_bin else 0 for b in range(5)]
    order = [4, 3, 2, 1, 0]
    assert list(B.r_slot_blk0) == list(np.concatenate([[0], np.cumsum([rb[b] for b in order])]))
    # combine items: tiles and big rows cover every blocked row once, inside their slice
    ct = view(B.ct_row, 2 * B.c_n_tiles).reshape(-1, 2).astype(np.int64)
    cs = view(B.ct_start, B.c_n_tiles * nb).reshape(-1, nb).astype(np.int64)
    cp = view(B.ct_pre, B.c_n_tiles * (nb + 1)).reshape(-1, nb + 1).astype(np.int64)
    ci = view(B.c_item, 2 * B.c_n_items).reshape(-1, 2).astype(np.int64)
    co = list(B.c_item_off)
    assert co[0] == 0 and co[-1] == B.c_n_items and co == sorted(co)
    rows_seen = np.zeros(n, np.int64)
    tiles_seen = []
    for x in range(NXCD):
        for first, kc in ci[co[x]:co[x + 1]]:
            if kc & 15 == ITEM_CBIG:
                assert sl[x] <= first < sl[x + 1]
                assert coff[first + 1] - coff[first] > prm.tile_nnz
                rows_seen[first] += 1
                continue
            assert kc & 15 == ITEM_CTILE
            tiles_seen.append(first)
            r0, r1 = ct[first]
            assert sl[x] <= r0 < r1 <= sl[x + 1] and r1 - r0 <= 4096
            rows_seen[r0:r1] += 1
            p0, npart = coff[r0], coff[r1] - coff[r0]
            assert npart <= prm.tile_nnz and cp[first, 0] == 0 and cp[first, nb] == npart
            got = []
            for b in range(nb):
                vs = np.arange(cs[first, b], cs[first, b] + cp[first, b + 1] - cp[first, b])
                assert np.all(vblk[vs] == b) and np.all((vrow[vs] >= r0) & (vrow[vs] < r1))
                got.extend(vs.tolist())
            want = cidx[coff[r0]:coff[r1]]
            assert sorted(got) == sorted(want.tolist())
            # lpos[p0 + k]: slot relative to the tile of the k-th partial the tile streams
            # (block order), so that row r's partials land in [coff[r] - p0, coff[r + 1] - p0)
            # in block order
            np.testing.assert_array_equal(want[lpos[p0:p0 + npart]], np.array(got))
    assert sorted(tiles_seen) == list(range(B.c_n_tiles))
    assert np.all(rows_seen[row0:] == 1) and np.all(rows_seen[:row0] == 0)


@pytest.mark.parametrize("hot", [0, 128], ids=["nohot", "hot"])
@pytest.mark.parametrize("min_bin", [0, 1, 3])
@pytest.mark.parametrize("small", [False, True], ids=["default", "small"])
def test_block_plan_invariants(min_bin, small, hot):
    g = ag.Graph.rmat(12)
    base = dict(SMALL) if small else {}
    prm = ag.plan_params(**dict(base, src_block_bits=8, blocked_min_bin=min_bin,
                                hot_segments=hot))
    st, p, err = build(g, prm)
    assert st == 0, err
    try:
        row0 = p.bin_row[min_bin]
        if row0 == g.n:   # no row in the blocked bins: nothing to block
            assert p.blk.bits == 0 and p.blk.nv == 0
            return
        assert p.blk.bits == 8 and p.blk.nblocks == 16 + p.blk.n_hot and p.blk.nv > 0
        assert (p.blk.n_hot > 0) == (hot > 0)
        check_plan(g, p)
        check_block_plan(g, p)
        if small:   # tiny tiles: some rows have more partials than a combine tile holds
            kinds = set(view(p.blk.c_item, 2 * p.blk.c_n_items)[1::2] & 15)
            assert {ITEM_CTILE, ITEM_CBIG} <= kinds
    finally:
        ag.lib().ag_plan_free(C.byref(p))


@pytest.mark.parametrize("piece", [1, 5])
def test_block_plan_piece_limit(piece):
    g = ag.Graph.rmat(12)
    st, p, err = build(g, ag.plan_params(src_block_bits=8, blocked_min_bin=0, piece_nnz=piece))
    assert st == 0, err
    try:
        assert p.blk.piece == piece
        if piece == 1:
            assert p.blk.nv == p.blk.mb
        check_plan(g, p)
        check_block_plan(g, p)
    finally:
        ag.lib().ag_plan_free(C.byref(p))


def test_block_plan_auto_piece_on_even_sources():
    """Sources spread evenly over the blocks (Erdos-Renyi, ~16 in-edges over 64 blocks):
    nearly one piece per in-edge, so the automatic limit makes every in-edge a piece; an RMAT
    graph of the same size keeps the default limit."""
    n = 1 << 14
    g = ag.Graph.erdos_renyi(n, 16 * n)
    st, p, err = build(g, ag.plan_params(src_block_bits=8, blocked_min_bin=0))
    assert st == 0, err
    try:
        assert p.blk.piece == 1 and p.blk.nv == p.blk.mb
        check_block_plan(g, p)
    finally:
        ag.lib().ag_plan_free(C.byref(p))
    st, p, err = build(ag.Graph.rmat(14), ag.plan_params(src_block_bits=8, blocked_min_bin=0))
    assert st == 0, err
    try:
        assert p.blk.piece == PIECE and p.blk.nv * 4 < p.blk.mb * 3
    finally:
        ag.lib().ag_plan_free(C.byref(p))


@pytest.mark.parametrize("sort", ["auto", "on", "off"])
@pytest.mark.parametrize("bits,piece", [(8, 0), (8, 1), (21, 0)])
def test_block_plan_tile_sort(sort, bits, piece):
    """Sorted pass-1 ids (tile_sort) on a skewed and an even graph, with multi-edge and
    single-edge pieces; blocks above 2^20 sources leave the tiles unsorted."""
    mode = {"auto": ag.SORT_AUTO, "on": ag.SORT_ON, "off": ag.SORT_OFF}[sort]
    for g in (ag.Graph.rmat(12), ag.Graph.erdos_renyi(1 << 12, 1 << 16)):
        st, p, err = build(g, ag.plan_params(src_block_bits=bits, blocked_min_bin=0,
                                             piece_nnz=piece, tile_sort=mode))
        assert st == 0, err
        try:
            assert p.blk.bits == bits
            assert p.blk.sorted == int(bits <= 20 and mode != ag.SORT_OFF)
            check_block_plan(g, p)
        finally:
            ag.lib().ag_plan_free(C.byref(p))


def test_block_plan_tile_sort_full_word():
    """2^20-source blocks and 4096-edge tiles fill the packed word: positions up to 4095 above
    20 bits of source offset (RMAT-21, 2 M vertices)."""
    g = ag.Graph.rmat(21)
    st, p, err = build(g, ag.plan_params(src_block_bits=20, blocked_min_bin=0))
    assert st == 0, err
    try:
        B = p.blk
        assert B.bits == 20 and B.sorted == 1
        it = view(B.item, 2 * B.n_items).reshape(-1, 2).astype(np.int64)
        check_sorted_tiles(p, it)
        vsp = view(B.vsp, B.mb).astype(np.int64)
        assert (vsp >> 20).max() == 4095 and (vsp & 0xFFFFF).max() == (1 << 20) - 1
    finally:
        ag.lib().ag_plan_free(C.byref(p))


def test_block_plan_long_pieces():
    # one row with 10000 in-edges from one source block is cut into pieces of 4096
    n = 10002
    src = np.concatenate([np.arange(1, 10001), [0, 1]]).astype(np.uint32)
    dst = np.concatenate([np.zeros(10000, np.uint32), [10001, 10001]]).astype(np.uint32)
    g = ag.Graph.from_edges(src, dst, n=n)
    st, p, err = build(g, ag.plan_params(src_block_bits=14, blocked_min_bin=0))
    assert st == 0, err
    try:
        check_block_plan(g, p)
        lens = np.diff(view(p.blk.voff, p.blk.nv + 1).astype(np.int64))
        assert sorted(lens.tolist())[-3:] == [1808, 4096, 4096]
        assert p.blk.c_n_tiles >= 1
    finally:
        ag.lib().ag_plan_free(C.byref(p))


def test_block_plan_auto_and_off():
    g = ag.Graph.rmat(10)
    for bits in (0, 255):   # auto leaves graphs of at most 2^21 vertices unblocked
        st, p, err = build(g, ag.plan_params(src_block_bits=bits))
        assert st == 0 and p.blk.bits == 0 and p.blk.nv == 0
        ag.lib().ag_plan_free(C.byref(p))
    st, p, err = build(g, ag.plan_params(src_block_bits=8, blocked_min_bin=5))
    assert st == 0 and p.blk.bits == 0
    ag.lib().ag_plan_free(C.byref(p))


def test_block_plan_rejects_bad_bits():
    g = ag.Graph.rmat(6)
    st, p, err = build(g, ag.plan_params(src_block_bits=7))
    assert st == 1 and b"src_block_bits" in err
    from accelgraph_amd import autotune
    cap = autotune.MAX_SRC_BLOCKS   # the autotuner's copy of the planner's limit
    n = cap * 256 + 1   # one block more than the limit, 256 sources each
    st, p, err = build(None, n=n, outdeg=np.zeros(n, np.uint32), idx=np.zeros(n, np.uint32),
                       deg=np.zeros(n, np.uint32), src=np.zeros(0, np.uint32),
                       params=ag.plan_params(src_block_bits=8))
    assert st == 1 and str(cap).encode() in err


@pytest.mark.parametrize("sort", [ag.SORT_OFF, ag.SORT_ON], ids=["unsorted", "sorted"])
def test_block_plan_weights_follow_sources(sort):
    """Weighted graph: the virtual-row weights are the row's weights regrouped exactly like its
    sources (SpMV reads vw[j] beside vsrc[j]); with sorted tiles they follow the packed ids
    (vw[j] beside vsp[j])."""
    g = ag.Graph.rmat(11)
    w = (np.arange(g.m, dtype=np.float64) + 1).astype(np.float32)   # unique per inverse edge
    st, p, err = build(g, ag.plan_params(src_block_bits=8, blocked_min_bin=0, tile_sort=sort),
                       weights=w)
    assert st == 0, err
    try:
        B, n = p.blk, g.n
        d2o = view(p.d2o, n)
        voff, vsrc = view(B.voff, B.nv + 1), view(B.vsrc, B.mb)
        vw = np.ctypeslib.as_array(B.vw, shape=(B.mb,)).copy()
        assert B.sorted == (sort == ag.SORT_ON)
        if B.sorted:   # back to virtual-row order through the packed positions
            vsp = view(B.vsp, B.mb).astype(np.int64)
            vr = np.full_like(vw, np.nan)
            for first, kc in view(B.item, 2 * B.n_items).reshape(-1, 2).astype(np.int64):
                assert kc & 15 == ITEM_VTILE
                e0, e1 = voff[first], voff[first + (kc >> 4)]
                vr[e0 + (vsp[e0:e1] >> B.bits)] = vw[e0:e1]
            vw = vr
        coff, cidx = view(B.coff, n + 1), view(B.cidx, B.nv)
        for r in range(B.row0, n):
            v = d2o[r]
            lo, hi = g.inv_idx[v], g.inv_idx[v] + g.in_degree[v]
            if hi == lo:
                assert coff[r + 1] == coff[r]
                continue
            osrc, ow = g.inv_src[lo:hi], w[lo:hi]
            dsrc = view(p.o2d, n)[osrc]
            order = np.argsort(block_keys(B, n)[dsrc], kind="stable")
            got_s = np.concatenate([vsrc[voff[q]:voff[q + 1]] for q in cidx[coff[r]:coff[r + 1]]])
            got_w = np.concatenate([vw[voff[q]:voff[q + 1]] for q in cidx[coff[r]:coff[r + 1]]])
            np.testing.assert_array_equal(got_s, dsrc[order])
            np.testing.assert_array_equal(got_w, ow[order])
    finally:
        ag.lib().ag_plan_free(C.byref(p))
