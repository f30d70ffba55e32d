"""Graph files (graph_io.c): text edge lists, binary edge lists, serialised CSR.

The reference's own test files pin the readers: 04_test_graphs/TEST/graphbrew/graph.org is
graph.bin written as text (tab- and space-separated), so both must load to the same arrays,
and graph.bin.txt is graph.bin relabelled and symmetrised, so it must load to a graph
isomorphic to the symmetrised graph.bin. The writers are pinned by round trips; the
serialised layout is this library's own and is pinned by its documented checksum."""
import os

import numpy as np
import pytest

import accelgraph_amd as ag

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def same_graph(a: ag.Graph, b: ag.Graph, weights: bool = True):
    assert (a.n, a.m) == (b.n, b.m)
    for name in ("out_degree", "edges_idx", "fwd_src", "fwd_dest", "in_degree", "inv_idx",
                 "inv_src"):
        np.testing.assert_array_equal(getattr(a, name), getattr(b, name), err_msg=name)
    for inv in (False, True):
        va, vb = a._v(inv), b._v(inv)
        np.testing.assert_array_equal(ag._view(va.in_degree, a.n, np.uint32),
                                      ag._view(vb.in_degree, b.n, np.uint32))
        ea, eb = a._e(inv), b._e(inv)
        np.testing.assert_array_equal(ag._view(ea.edges_array_src, a.m, np.uint32),
                                      ag._view(eb.edges_array_src, b.m, np.uint32))
    if weights:
        assert (a.fwd_weight is None) == (b.fwd_weight is None)
        if a.fwd_weight is not None:
            np.testing.assert_array_equal(a.fwd_weight, b.fwd_weight)
            np.testing.assert_array_equal(a.inv_weight, b.inv_weight)
    ga, gb = a.ptr.contents, b.ptr.contents
    assert ga.avg_degree == gb.avg_degree
    if weights:
        assert ga.max_weight == gb.max_weight


def write(tmp_path, name, text, mode="w"):
    p = tmp_path / name
    with open(p, mode, newline="") as f:
        f.write(text)
    return str(p)


# ------------------------------------------------------------------ reference files

@pytest.mark.parametrize("simplify", [True, False])
def test_reference_text_copy_equals_binary(simplify):
    """graph.org holds graph.bin's 19 edges as text ("1\\t2", "1   7", ...)."""
    a = ag.Graph.load_text(os.path.join(GOLDEN, "graphbrew_graph.org"), simplify=simplify)
    b = ag.Graph.load_binary(os.path.join(GOLDEN, "graphbrew_graph.bin"), simplify=simplify)
    assert (a.n, a.m) == (13, 19)
    same_graph(a, b)
    c = ag.Graph.load_file(os.path.join(GOLDEN, "graphbrew_graph.org"),
                           ag.FILE_TEXT_EDGE_LIST, simplify)
    same_graph(a, c)


def _isomorphism(adj_a, adj_b):
    """Backtracking search for a vertex map p with p(adj_a) == adj_b (small graphs)."""
    n = len(adj_a)
    deg_a = [len(x) for x in adj_a]
    deg_b = [len(x) for x in adj_b]
    if sorted(deg_a) != sorted(deg_b):
        return None
    order = sorted(range(n), key=lambda v: -deg_a[v])
    p, used = [-1] * n, [False] * n

    def go(k):
        if k == n:
            return True
        v = order[k]
        for w in range(n):
            if used[w] or deg_b[w] != deg_a[v]:
                continue
            ok = all((p[u] in adj_b[w]) == (u in adj_a[v]) for u in order[:k])
            if ok:
                p[v], used[w] = w, True
                if go(k + 1):
                    return True
                p[v], used[w] = -1, False
        return False

    return p if go(0) else None


def test_reference_symmetrised_relabelled_copy():
    """graph.bin.txt: 38 lines, graph.bin with each edge in both directions under a
    relabelling of the vertices."""
    t = ag.Graph.load_text(os.path.join(GOLDEN, "graphbrew_graph.bin.txt"))
    assert (t.n, t.m) == (13, 38)
    pairs = set(zip(t.fwd_src.tolist(), t.fwd_dest.tolist()))
    assert all((d, s) in pairs for s, d in pairs)                       # symmetric
    np.testing.assert_array_equal(t.out_degree, t.in_degree)
    raw = np.fromfile(os.path.join(GOLDEN, "graphbrew_graph.bin"), dtype="<u4").reshape(-1, 2)
    adj_bin = [set() for _ in range(13)]
    for s, d in raw.tolist():
        adj_bin[s].add(d)
        adj_bin[d].add(s)
    adj_txt = [set() for _ in range(13)]
    for s, d in pairs:
        adj_txt[s].add(d)
    p = _isomorphism(adj_bin, adj_txt)
    assert p is not None and sorted(p) == list(range(13))
    assert {(p[s], p[d]) for s, d in raw.tolist()} <= pairs


# ------------------------------------------------------------------ text syntax

def test_text_syntax_variants(tmp_path):
    text = ("# comment line\n"
            "% another comment\n"
            "\n"
            "0 1\n"
            "1\t2\r\n"
            "  2,3  \n"
            "3 \t 0 extra columns 99\n"
            "\t\n"
            "4    4\n"
            "2 3\n"
            "5 0")                                   # no final newline
    g = ag.Graph.load_text(write(tmp_path, "v.txt", text))
    ref = ag.Graph.from_edges([0, 1, 2, 3, 4, 2, 5], [1, 2, 3, 0, 4, 3, 0])
    same_graph(g, ref)
    g = ag.Graph.load_text(write(tmp_path, "v.txt", text), simplify=False)
    same_graph(g, ag.Graph.from_edges([0, 1, 2, 3, 4, 2, 5], [1, 2, 3, 0, 4, 3, 0],
                                      simplify=False))


def test_text_weights(tmp_path):
    text = "0 1 0.5\n1 2 1e-3\n2 0 -2.25\n0 1 7\n3 3 1\n"
    g = ag.Graph.load_text(write(tmp_path, "w.txt", text), weighted=True)
    ref = ag.Graph.from_edges([0, 1, 2, 0, 3], [1, 2, 0, 1, 3],
                              weights=np.float32([0.5, 1e-3, -2.25, 7, 1]))
    same_graph(g, ref)
    assert list(g.fwd_weight) == [np.float32(0.5), np.float32(1e-3), np.float32(-2.25)]
    # unweighted read ignores the third column
    same_graph(ag.Graph.load_text(write(tmp_path, "w.txt", text)),
               ag.Graph.from_edges([0, 1, 2, 0, 3], [1, 2, 0, 1, 3]))


@pytest.mark.parametrize("text,weighted,line", [
    ("0 1\n1\n", False, 2),                 # missing destination
    ("0 1\n-1 2\n", False, 2),              # negative id
    ("0 1\n\n# c\n1 x\n", False, 4),        # not a number
    ("0 1\n1 2x\n", False, 2),              # junk glued to an id
    ("0 4294967295\n", False, 1),           # id == 2^32-1 (no room for the vertex count)
    ("0 99999999999\n", False, 1),          # overflow
    ("0 1 2\n1 2\n", True, 2),              # weighted, weight missing
    ("0 1 nan\n", True, 1),                 # non-finite weight
    ("0 1 1.5q\n", True, 1),                # junk weight
])
def test_text_errors_name_the_line(tmp_path, text, weighted, line):
    path = write(tmp_path, "bad.txt", text)
    with pytest.raises(ag.AGError) as e:
        ag.Graph.load_text(path, weighted=weighted)
    assert f"bad.txt:{line}:" in str(e.value)


def test_text_empty_and_comment_only(tmp_path):
    for text in ("", "# nothing\n\n% here\n"):
        g = ag.Graph.load_text(write(tmp_path, "e.txt", text))
        assert (g.n, g.m) == (0, 0)


def test_text_missing_file(tmp_path):
    with pytest.raises(ag.AGError, match="cannot open"):
        ag.Graph.load_text(str(tmp_path / "absent.txt"))


def test_text_parallel_chunks_keep_file_order(tmp_path):
    """A file of several MiB is split over threads at line starts. Duplicate edges with
    different weights make the file order visible: the first occurrence keeps its weight."""
    rng = np.random.default_rng(3)
    m = 400000
    src = rng.integers(0, 20000, m).astype(np.uint32)
    dst = rng.integers(0, 20000, m).astype(np.uint32)
    src[1::3], dst[1::3] = src[::3][:src[1::3].size], dst[::3][:dst[1::3].size]
    w = rng.random(m).astype(np.float32)
    lines = [f"{s}\t{d}\t{x!r}" for s, d, x in zip(src.tolist(), dst.tolist(), w.tolist())]
    for i in range(0, m, 997):                       # comments and blank lines in between
        lines[i] = "# c\n\n" + lines[i]
    path = write(tmp_path, "big.txt", "\n".join(lines) + "\n")
    assert os.path.getsize(path) > (4 << 20)
    g = ag.Graph.load_text(path, weighted=True)
    same_graph(g, ag.Graph.from_edges(src, dst, weights=w))
    g = ag.Graph.load_text(path, weighted=True, simplify=False)
    same_graph(g, ag.Graph.from_edges(src, dst, weights=w, simplify=False))


# ------------------------------------------------------------------ writers and round trips

@pytest.mark.parametrize("weighted", [False, True])
def test_round_trips(tmp_path, weighted):
    g = ag.Graph.rmat(14)
    if weighted:
        g.set_weights(np.random.default_rng(5).random(g.m).astype(np.float32) * 10 - 3)
    # edge lists carry no vertex count: trailing isolated vertices (RMAT-14 has some) are
    # not in the file, so the loaded graph is the one rebuilt from the edges alone
    assert g.out_degree[-1] == 0 and g.in_degree[-1] == 0
    e = ag.Graph.from_edges(g.fwd_src, g.fwd_dest, weights=g.fwd_weight)
    assert e.n < g.n and e.m == g.m
    np.testing.assert_array_equal(e.fwd_dest, g.fwd_dest)
    p = str(tmp_path / "g")
    g.save(p + ".txt", ag.FILE_TEXT_EDGE_LIST)
    same_graph(ag.Graph.load_text(p + ".txt", weighted=weighted), e, weights=weighted)
    g.save(p + ".bin", ag.FILE_BINARY_EDGE_LIST)
    assert os.path.getsize(p + ".bin") == 8 * g.m
    same_graph(ag.Graph.load_binary(p + ".bin"), e, weights=False)
    g.save(p + ".csr")
    same_graph(ag.Graph.load_serialized(p + ".csr"), g)
    same_graph(ag.Graph.load_file(p + ".csr", ag.FILE_CSR_SERIALIZED), g)


def test_serialized_keeps_unsimplified_and_empty(tmp_path):
    g = ag.Graph.from_edges([0, 0, 1, 2, 2], [1, 1, 1, 0, 2], n=6, simplify=False)
    g.save(str(tmp_path / "u.csr"))
    same_graph(ag.Graph.load_serialized(str(tmp_path / "u.csr")), g)
    e = ag.Graph.from_edges([], [], n=4)
    e.save(str(tmp_path / "e.csr"))
    same_graph(ag.Graph.load_serialized(str(tmp_path / "e.csr")), e)


def checksum(words: np.ndarray) -> int:
    """The file's trailer: sum of w_j * (2j + 1) over its u32 words, modulo 2^64."""
    j = np.arange(words.size, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return int(np.sum(words.astype(np.uint64) * (2 * j + 1), dtype=np.uint64))


def test_serialized_layout_and_checksum(tmp_path):
    g = ag.Graph.rmat(10)
    path = str(tmp_path / "g.csr")
    g.save(path)
    raw = np.fromfile(path, dtype=np.uint8)
    n, m = g.n, g.m
    assert raw.size == 40 + 4 * (6 * n + 4 * m) + 8
    assert bytes(raw[:8]) == b"AGCSR\0\0\0"
    words = raw[:-8].view("<u4")
    assert list(words[2:6]) == [1, 0, n, m]
    assert checksum(words) == int(raw[-8:].view("<u8")[0])
    # forward out_degree, in_degree, edges_idx, src, dest in order after the header
    np.testing.assert_array_equal(words[10:10 + n], g.out_degree)
    np.testing.assert_array_equal(words[10 + 2 * n:10 + 3 * n], g.edges_idx)
    np.testing.assert_array_equal(words[10 + 3 * n + m:10 + 3 * n + 2 * m], g.fwd_dest)


def test_serialized_damage_is_rejected(tmp_path):
    g = ag.Graph.rmat(10)
    path = str(tmp_path / "g.csr")
    g.save(path)
    good = np.fromfile(path, dtype=np.uint8)
    rng = np.random.default_rng(11)
    for off in [40, 41, 40 + 4 * g.n + 3, good.size - 9, good.size - 1,
                *rng.integers(40, good.size, 8).tolist()]:
        bad = good.copy()
        bad[off] ^= np.uint8(1 << int(rng.integers(0, 8)))
        bad.tofile(path)
        with pytest.raises(ag.AGError, match="checksum"):
            ag.Graph.load_serialized(path)
    good[:-4].tofile(path)
    with pytest.raises(ag.AGError, match="size"):
        ag.Graph.load_serialized(path)
    bad = good.copy()
    bad[0] = ord("X")
    bad.tofile(path)
    with pytest.raises(ag.AGError, match="not a serialised CSR"):
        ag.Graph.load_serialized(path)


@pytest.mark.parametrize("what", ["dest_range", "offset", "row_id", "unsorted", "inverse"])
def test_serialized_structure_is_checked(tmp_path, what):
    """Files with a valid checksum whose arrays would send the kernels out of bounds or
    disagree between the two views."""
    g = ag.Graph.from_edges([0, 0, 1, 2, 3], [1, 2, 2, 3, 0], n=5)
    path = str(tmp_path / "g.csr")
    g.save(path)
    raw = np.fromfile(path, dtype=np.uint8)
    words = raw[:-8].view("<u4").copy()
    n, m = g.n, g.m
    fwd_idx, fwd_src, fwd_dst = 10 + 2 * n, 10 + 3 * n, 10 + 3 * n + m
    inv_dst = 10 + 3 * n + 2 * m + 3 * n + m
    if what == "dest_range":
        words[fwd_dst + 1] = n
    elif what == "offset":
        words[fwd_idx + 1] += 1
    elif what == "row_id":
        words[fwd_src + 2] = 4
    elif what == "unsorted":
        words[fwd_dst], words[fwd_dst + 1] = words[fwd_dst + 1], words[fwd_dst]
    else:                                        # inverse row points elsewhere
        words[fwd_dst + 4] = 1                   # 3 -> 0 becomes 3 -> 1 in the forward view
    out = np.concatenate([words.view(np.uint8),
                          np.array([checksum(words)], dtype="<u8").view(np.uint8)])
    out.tofile(path)
    with pytest.raises(ag.AGError, match="inconsistent"):
        ag.Graph.load_serialized(path)


def test_file_format_codes(tmp_path):
    with pytest.raises(ag.AGError, match="file format 3"):
        ag.Graph.load_file(os.path.join(GOLDEN, "graphbrew_graph.bin"), 3)
    g = ag.Graph.load_file(os.path.join(GOLDEN, "graphbrew_graph.bin"), ag.FILE_BINARY_EDGE_LIST)
    assert (g.n, g.m) == (13, 19)
    with pytest.raises(ag.AGError):
        g.save(str(tmp_path / "x"), 7)
