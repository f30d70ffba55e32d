"""Host C code under AddressSanitizer and UndefinedBehaviorSanitizer: the graph builders and
file readers (fixtures, malformed text lines, every truncation and a sweep of corrupted bytes
of a serialised CSR) and the planner (every plan shape of the parity tests). Built with gcc
from the library's own host sources; no GPU code is involved."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "accelgraph_amd", "host")
HERE = os.path.join(ROOT, "tests", "sanitize")
FLAGS = ["-O1", "-g", "-fopenmp", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-fno-omit-frame-pointer", f"-I{ROOT}/include", f"-I{HOST}"]


def build_and_run(tmp_path, driver, sources, args=()):
    exe = tmp_path / driver.replace(".c", "")
    cmd = ["gcc", *FLAGS, os.path.join(HERE, driver), *[os.path.join(HOST, s) for s in sources],
           "-o", str(exe), "-lm"]
    subprocess.run(cmd, check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")
    res = subprocess.run([str(exe), *args], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True, env=env, timeout=600)
    assert res.returncode == 0, res.stdout[-4000:]
    assert "failures 0" in res.stdout


def test_graph_readers_under_sanitizers(tmp_path):
    build_and_run(tmp_path, "graph_io_driver.c", ["graph_io.c", "graph_csr.c", "mt19937.c"],
                  (os.path.join(ROOT, "tests", "golden"), str(tmp_path)))


def test_planner_under_sanitizers(tmp_path):
    build_and_run(tmp_path, "plan_driver.c", ["plan.c", "graph_csr.c", "mt19937.c"])
