"""Graph construction on the GPU against the host builders (graph_csr.c): the chunked MT19937
edge streams bit for bit, and every array of the resulting struct GraphCSR."""
import ctypes as C
import sys

import numpy as np
import pytest

import accelgraph_amd as ag

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_graph(a: ag.Graph, b: ag.Graph):
    assert (a.n, a.m) == (b.n, b.m)
    for name in ("out_degree", "edges_idx", "fwd_src", "fwd_dest", "in_degree", "inv_idx",
                 "inv_src"):
        np.testing.assert_array_equal(getattr(a, name), getattr(b, name), err_msg=name)
    # the remaining arrays of both views
    for inv in (False, True):
        va, vb = a._v(inv), b._v(inv)
        np.testing.assert_array_equal(ag._view(va.in_degree, a.n, np.uint32),
                                      ag._view(vb.in_degree, b.n, np.uint32))
        ea, eb = a._e(inv), b._e(inv)
        np.testing.assert_array_equal(ag._view(ea.edges_array_src, a.m, np.uint32),
                                      ag._view(eb.edges_array_src, b.m, np.uint32))
        assert (ea.num_edges, ea.num_vertices) == (eb.num_edges, eb.num_vertices)
    if a.fwd_weight is not None or b.fwd_weight is not None:
        np.testing.assert_array_equal(a.fwd_weight, b.fwd_weight)
        np.testing.assert_array_equal(a.inv_weight, b.inv_weight)
    ga, gb = a.ptr.contents, b.ptr.contents
    assert (ga.max_weight, ga.avg_degree) == (gb.max_weight, gb.avg_degree)


@pytest.mark.parametrize("scale,m", [(10, 16 << 10), (12, 3 * (1 << 20) + 12345), (21, 16 << 21),
                                     (31, 5000)])
def test_rmat_edge_stream_matches_host(device, scale, m):
    """Chunked MT19937 streams: one chunk, several chunks with a partial last one, 32 chunks,
    and the deepest recursion (31 draws per edge, the most any edge spans)."""
    say(f"rmat edges scale={scale} m={m}")
    L = ag.lib()
    u32p = C.POINTER(C.c_uint32)
    hs, hd = np.zeros(m, np.uint32), np.zeros(m, np.uint32)
    assert L.graphGenerateRMATEdges(scale, m, 0.57, 0.19, 0.19, ag.DEFAULT_SEED,
                                    hs.ctypes.data_as(u32p), hd.ctypes.data_as(u32p)) == 0
    gs, gd = np.zeros(m, np.uint32), np.zeros(m, np.uint32)
    assert L.ag_generate_rmat_edges(device.handle, scale, m, 0.57, 0.19, 0.19, ag.DEFAULT_SEED,
                                    gs.ctypes.data_as(u32p), gd.ctypes.data_as(u32p)) == 0
    np.testing.assert_array_equal(gs, hs)
    np.testing.assert_array_equal(gd, hd)


@pytest.mark.parametrize("simplify", [True, False])
def test_rmat_csr_matches_host(device, simplify):
    say(f"rmat csr simplify={simplify}")
    same_graph(ag.Graph.rmat_gpu(device, 14, simplify=simplify),
               ag.Graph.rmat(14, simplify=simplify))


def test_rmat20_csr_matches_host(device):
    say("rmat-20 csr")
    same_graph(ag.Graph.rmat_gpu(device, 20), ag.Graph.rmat(20))


@pytest.mark.parametrize("n,m", [(1 << 12, 16 << 12), (1000, 3 * (1 << 20) + 7), (37, 5000)])
def test_er_csr_matches_host(device, n, m):
    say(f"er csr n={n} m={m}")
    same_graph(ag.Graph.erdos_renyi_gpu(device, n, m), ag.Graph.erdos_renyi(n, m))


@pytest.mark.parametrize("simplify", [True, False])
def test_edge_list_with_duplicates_loops_weights(device, simplify):
    """Duplicates (the first occurrence keeps its weight), self-loops, isolated vertices,
    an explicit vertex count and the inferred one."""
    rng = np.random.default_rng(7)
    m = 200000
    src = rng.integers(0, 3000, m).astype(np.uint32)
    dst = rng.integers(0, 3000, m).astype(np.uint32)
    src[::17] = dst[::17]                              # self-loops
    src[1::5], dst[1::5] = src[::5][:src[1::5].size], dst[::5][:dst[1::5].size]  # duplicates
    w = rng.random(m).astype(np.float32)
    for n in (0, 5000):
        say(f"edge list simplify={simplify} n={n}")
        same_graph(ag.Graph.from_edges_gpu(device, src, dst, n=n, weights=w, simplify=simplify),
                   ag.Graph.from_edges(src, dst, n=n, weights=w, simplify=simplify))
        same_graph(ag.Graph.from_edges_gpu(device, src, dst, n=n, simplify=simplify),
                   ag.Graph.from_edges(src, dst, n=n, simplify=simplify))


def test_edge_list_corner_cases(device):
    say("edge list corner cases")
    same_graph(ag.Graph.from_edges_gpu(device, [], [], n=5), ag.Graph.from_edges([], [], n=5))
    same_graph(ag.Graph.from_edges_gpu(device, [3], [3], n=0), ag.Graph.from_edges([3], [3], n=0))
    same_graph(ag.Graph.from_edges_gpu(device, [0, 1, 2], [1, 2, 0]),
               ag.Graph.from_edges([0, 1, 2], [1, 2, 0]))
    with pytest.raises(ag.AGError):
        ag.Graph.from_edges_gpu(device, [0, 9], [1, 2], n=5)     # id >= num_vertices


def test_gpu_built_graph_runs_the_path(device):
    """A GPU-built graph is an ordinary struct GraphCSR: PageRank on it equals PageRank on the
    host-built one bit for bit (fixed point) and value for value (FP32)."""
    say("gpu-built graph through pagerank")
    a, b = ag.Graph.rmat_gpu(device, 16), ag.Graph.rmat(16)
    da, db = device.upload(a), device.upload(b)
    ra = da.pagerank(kind=32, max_iterations=20)
    rb = db.pagerank(kind=32, max_iterations=20)
    np.testing.assert_array_equal(ra.ranks.view(np.uint32), rb.ranks.view(np.uint32))
    da.close()
    db.close()
