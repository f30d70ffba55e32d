"""The drop-in host code without a GPU: the reference's entry points and the graph cache
(accelgraph_amd/host/pageRank.c, SPMV.c, accel_context.c) linked against a CPU fake of the
device layer (tests/fake/ag_fake.c), the way the reference tests its host runtime against a
fake libcxl (02_capi_graph test_capienv_verification.c:78-299). A run without faults must
reach the oracle's results through the same calls; each forced failure (device open, upload,
session start, a failed kernel, an expired wait, the history and result copies, the SpMV)
must print one line naming the call and its status and exit with EXIT_FAILURE, as the
reference's runtime does on every accelerator failure (capienv.c:80-132)."""
import os
import subprocess

import numpy as np
import pytest

import accelgraph_amd as ag
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "accelgraph_amd", "host")
SOURCES = [os.path.join(ROOT, "tests", "fake", "driver.c"),
           os.path.join(ROOT, "tests", "fake", "ag_fake.c"),
           os.path.join(ROOT, "oracle", "ag_oracle.c")] + \
          [os.path.join(HOST, f) for f in ("pageRank.c", "SPMV.c", "accel_context.c", "plan.c",
                                           "graph_csr.c", "mt19937.c")]


@pytest.fixture(scope="module")
def fake_exe(tmp_path_factory):
    exe = tmp_path_factory.mktemp("fake") / "fake_driver"
    subprocess.run(["gcc", "-O1", "-fopenmp", f"-I{os.path.join(ROOT, 'include')}", f"-I{HOST}",
                    *SOURCES, "-lm", "-o", str(exe)], check=True)
    return exe


def run(exe, fail=None):
    env = dict(os.environ, OMP_NUM_THREADS="2")
    env.pop("AG_FAKE_FAIL", None)
    if fail:
        env["AG_FAKE_FAIL"] = fail
    return subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=120, env=env)


def test_fake_backend_happy_path(fake_exe):
    res = run(fake_exe)
    assert res.returncode == 0, res.stderr
    lines = res.stdout.splitlines()
    g = ag.Graph.rmat(10)
    for code, kind in ((0, 0), (9, 32)):
        n, iterations = (int(x) for x in
                         next(l for l in lines if l.startswith(f"PR {code} ")).split()[2:])
        ranks = np.array([int(l.split()[1], 16) for l in lines if l.startswith(f"R{code} ")],
                         np.uint32)
        ref = oracle.pagerank_pull(g.out_degree, g.inv_idx, g.in_degree, g.inv_src, kind=kind,
                                   max_iterations=100, epsilon=1e-4)
        assert n == g.n and ranks.size == n
        assert iterations == ref.iterations
        np.testing.assert_array_equal(ranks, ref.ranks.view(np.uint32))
    # the reference's iteration table and summary rows were printed
    assert sum("| Iterations | PR Sum" in l for l in lines) == 2
    x = np.where(g.out_degree > 0, np.float32(1) / np.maximum(g.out_degree, 1).astype(np.float32),
                 np.float32(0.001)).astype(np.float32)
    y = np.array([int(l.split()[1], 16) for l in lines if l.startswith("Y ")], np.uint32)
    y_ref = oracle.spmv_f32(g.inv_idx, g.in_degree, g.inv_src, None, x)
    np.testing.assert_array_equal(y, y_ref.view(np.uint32))


@pytest.mark.parametrize("fail,call,status", [
    ("open", "ag_device_open", "AG_ERR_NO_DEVICE"),
    ("upload", "ag_graph_upload", "AG_ERR_OOM"),
    ("begin", "ag_pr_begin", "AG_ERR_HIP"),
    ("device", "ag_pr_step", "AG_ERR_HIP"),
    ("timeout", "ag_pr_step", "AG_ERR_TIMEOUT"),
    ("history", "ag_pr_history", "AG_ERR_HIP"),
    ("end", "ag_pr_end", "AG_ERR_HIP"),
    ("spmv", "ag_spmv_pull", "AG_ERR_TIMEOUT"),
])
def test_fake_backend_failures_exit(fake_exe, fail, call, status):
    res = run(fake_exe, fail)
    assert res.returncode == 1
    err = [l for l in res.stderr.splitlines() if l.startswith("[accelgraph]")]
    assert len(err) == 1 and err[0].startswith(f"[accelgraph] {call} failed: {status} (fake: ")
    # nothing after the failing call ran: no summary row of a later stage
    if fail != "spmv":
        assert "Starting SPMV-PULL" not in res.stdout
