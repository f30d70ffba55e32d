"""A plain C caller of the drop-in entry points, as an AccelGraph maintainer would write one:
compiled with gcc against include/ and linked with -laccelgraph, no Python in between. The
CPU test builds and links it; the GPU test runs it (fixed-32 PageRank and FP32 SpMV through
pageRankGraphCSR / SPMVGraphCSR on an RMAT-10 graph built by graphCSRGenerateRMAT) and checks
its output against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import accelgraph_amd as ag
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.dirname(ag._abi.LIB_PATH)

PROGRAM = r"""
#include <stdio.h>
#include <string.h>
#include "graphCSR.h"
#include "pageRank.h"
#include "SPMV.h"
#include "accelgraph_compat.h"

int main(void)
{
    struct GraphCSR *g = graphCSRGenerateRMAT(10, 16, 0.57, 0.19, 0.19, 27491095u, 1);
    if (!g)
        return 2;
    struct Arguments a;
    memset(&a, 0, sizeof(a));
    a.iterations = 100;
    a.epsilon = 1e-4;
    a.pushpull = 9;                      /* fixed-point 32 pull */
    a.ker_numThreads = 1;
    a.afu_config = a.afu_config_2 = a.cu_config = a.cu_config_2 = 1;
    struct PageRankStats *st = pageRankGraphCSR(&a, g);
    printf("N %u %u\n", g->num_vertices, st->iterations);
    for (uint32_t v = 0; v < g->num_vertices; v++) {
        uint32_t bits;
        memcpy(&bits, &st->pageRanks[v], 4);
        printf("R %08x\n", bits);
    }
    freePageRankStats(st);
    a.pushpull = 0;                      /* FP32 SpMV pull, 1 repetition */
    a.iterations = 1;
    struct SPMVStats *sp = SPMVGraphCSR(&a, g);
    for (uint32_t v = 0; v < g->num_vertices; v++)
        printf("Y %.9g\n", sp->vector_output[v]);
    freeSPMVStats(sp);
    accelgraphShutdown();
    graphCSRFree(g);
    return 0;
}
"""


def build_program(tmp_path):
    c = tmp_path / "caller.c"
    c.write_text(PROGRAM)
    exe = tmp_path / "caller"
    subprocess.run(["gcc", "-O2", f"-I{INCLUDE}", str(c), "-o", str(exe), f"-L{LIBDIR}",
                    "-laccelgraph", f"-Wl,-rpath,{LIBDIR}"], check=True)
    return exe


def test_c_caller_builds_and_links(tmp_path):
    exe = build_program(tmp_path)
    assert exe.exists()
    # every symbol the program needs resolves against the library
    out = subprocess.run(["ldd", str(exe)], stdout=subprocess.PIPE, text=True, check=True).stdout
    assert "libaccelgraph.so" in out and "not found" not in out


@pytest.mark.gpu
def test_c_caller_runs(tmp_path):
    exe = build_program(tmp_path)
    print("C caller: fixed-32 PageRank and FP32 SpMV on RMAT-10", flush=True)
    res = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, timeout=120)
    assert res.returncode == 0
    lines = res.stdout.splitlines()
    n, iterations = (int(x) for x in next(l for l in lines if l.startswith("N ")).split()[1:])
    ranks = np.array([int(l.split()[1], 16) for l in lines if l.startswith("R ")], np.uint32)
    y = np.array([float(l.split()[1]) for l in lines if l.startswith("Y ")], np.float32)
    g = ag.Graph.rmat(10)
    assert n == g.n and ranks.size == n and y.size == n
    ref = oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src, kind=32,
                               max_iterations=100, epsilon=1e-4)
    assert iterations == ref.iterations
    np.testing.assert_array_equal(ranks, ref.ranks.view(np.uint32))
    # SPMV.c:694-701 x (1/out-degree, 0.001 for sinks), unweighted: constant weight 0.0001f
    x = np.where(g.out_degree > 0, np.float32(1) / np.maximum(g.out_degree, 1).astype(np.float32),
                 np.float32(0.001)).astype(np.float32)
    y_ref = oracle.spmv_f32(g.inv_idx, g.in_degree, g.inv_src, None, x)
    np.testing.assert_allclose(y, y_ref, rtol=1e-6, atol=0)


UNSUPPORTED = r"""
#include <stdlib.h>
#include <string.h>
#include "graphCSR.h"
#include "pageRank.h"

int main(int argc, char **argv)
{
    uint32_t src[2] = {0, 1}, dst[2] = {1, 0};
    struct GraphCSR *g = graphCSRFromEdgeList(2, 2, src, dst, NULL, 1);
    struct Arguments a;
    memset(&a, 0, sizeof(a));
    a.iterations = 10;
    a.epsilon = 1e-4;
    a.pushpull = (uint32_t)atoi(argv[1]);
    pageRankGraphCSR(&a, g);   /* exits for the codes this library does not provide */
    return 0;
}
"""


@pytest.mark.parametrize("code,what", [(5, "quantised push"), (7, "data-driven push"),
                                       (8, "data-driven pull-push")])
def test_unprovided_codes_exit_with_a_message(tmp_path, code, what):
    """pageRankGraphCSR codes 5, 7 and 8 print one line naming the variant and exit with
    EXIT_FAILURE, the reference's convention for a failed accelerator path; nothing touches the
    device before that."""
    c = tmp_path / "unsupported.c"
    c.write_text(UNSUPPORTED)
    exe = tmp_path / "unsupported"
    subprocess.run(["gcc", "-O2", f"-I{INCLUDE}", str(c), "-o", str(exe), f"-L{LIBDIR}",
                    "-laccelgraph", f"-Wl,-rpath,{LIBDIR}"], check=True)
    res = subprocess.run([str(exe), str(code)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                         text=True, timeout=60, env=dict(os.environ, HIP_VISIBLE_DEVICES=""))
    assert res.returncode == 1
    assert f"pushpull={code} ({what}) is not provided" in res.stderr
