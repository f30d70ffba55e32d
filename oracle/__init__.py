"""CPU oracle: ctypes wrapper of oracle/liboracle.so (a C restatement of the reference's
PageRank-pull / SpMV-pull path; see ag_oracle.c for the file:line it follows).

TEST INFRASTRUCTURE. Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU leg
(the CPU baseline timings and the FP64 reference its error figures are measured against), and
only as the checker / CPU baseline. The product (accelgraph_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

u32, u64, f32, f64 = C.c_uint32, C.c_uint64, C.c_float, C.c_double
P = C.POINTER


class _PrOut(C.Structure):
    _fields_ = [("iterations", u32), ("executed", u32), ("active", u32), ("error_total", f64),
                ("rank_sum", f64)]


_LIB = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_fixed_encode.restype = u64
        L.or_fixed_encode.argtypes = [u32, f32]
        L.or_fixed_decode.restype = f64
        L.or_fixed_decode.argtypes = [u32, u64]
        L.or_gather_f32.argtypes = [u32, P(u32), P(u32), P(u32), P(f32), P(f32)]
        L.or_gather_u64.argtypes = [u32, P(u32), P(u32), P(u32), P(u64), P(u64)]
        L.or_spmv_f32.argtypes = [u32, P(u32), P(u32), P(u32), P(f32), P(f32), P(f32), C.c_int]
        L.or_spmv_fixed.argtypes = [u32, P(u32), P(u32), P(u32), P(u32), u32, P(u32), P(u64),
                                    C.c_int]
        L.or_pagerank_pull.restype = C.c_int
        L.or_pagerank_pull.argtypes = [u32, P(u32), P(u32), P(u32), P(u32), u32, u32, f64, u32,
                                       f64, u32, P(f32), C.c_int, P(f32), P(f32), P(u32), P(f64),
                                       P(_PrOut)]
        L.or_pagerank_dd_pull.restype = C.c_int
        L.or_pagerank_dd_pull.argtypes = [u32, P(u32), P(u32), P(u32), P(u32), P(u32), P(u32),
                                          P(u32), u32, f64, u32, P(f32), C.c_int, P(f32),
                                          P(f32), P(u32), P(f64), P(_PrOut)]
        L.or_pagerank_push.restype = C.c_int
        L.or_pagerank_push.argtypes = [u32, P(u32), P(u32), P(u32), P(u32), u32, u32, f64, u32,
                                       f64, u32, P(f32), C.c_int, P(f32), P(f32), P(u32),
                                       P(f64), P(_PrOut)]
        L.or_spmv_push_f32.argtypes = [u32, P(u32), P(u32), P(u32), P(f32), P(f32), u32,
                                       C.c_int, P(f32)]
        L.or_spmv_push_fixed.argtypes = [u32, P(u32), P(u32), P(u32), P(f32), P(u64), u32,
                                         C.c_int, P(u64)]
        L.or_mul_fixed64_check.restype = u64
        L.or_mul_fixed64_check.argtypes = [u64, u64]
        L.or_quantize.restype = f32
        L.or_quantize.argtypes = [u32, P(f32), u32, P(u64)]
        L.or_pagerank_f64.restype = C.c_int
        L.or_pagerank_f64.argtypes = [u32, P(u32), P(u32), P(u32), P(u32), f64, u32, C.c_int,
                                      P(f64)]
        L.or_ppr_f64.restype = C.c_int
        L.or_ppr_f64.argtypes = [u32, P(u32), P(u32), P(u32), P(u32), u32, f64, u32, C.c_int,
                                 P(f64)]
        L.or_fixed_encode_n.argtypes = [u32, P(f32), u64, P(u64)]
        L.or_ppr_residual.restype = C.c_int
        L.or_ppr_residual.argtypes = [u32, P(u32), P(u32), P(u32), P(u32), u32, P(u32), P(f32),
                                      C.c_int, P(f64), P(f64)]
        L.or_build_csr.restype = C.c_int
        L.or_build_csr.argtypes = [u32, u64, P(u32), P(u32), P(u32), P(u32), P(u32), P(u32),
                                   P(u32), P(u32), P(u64)]
        L.or_max_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _p(a, t):
    return None if a is None else a.ctypes.data_as(P(t))


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def fixed_encode(bits: int, x) -> np.ndarray:
    L = lib()
    return np.array([L.or_fixed_encode(bits, float(v)) for v in np.asarray(x, np.float32)],
                    dtype=np.uint64)


def fixed_encode_array(bits: int, x, threads: int = 0) -> np.ndarray:
    """fixed_encode over a whole array in C (OpenMP): for arrays of millions of values."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(max(x.size, 1), np.uint64)
    lib().or_fixed_encode_n(bits, _p(x, f32), x.size, _p(out, u64))
    return out[:x.size]


def fixed_decode(bits: int, q) -> np.ndarray:
    L = lib()
    return np.array([L.or_fixed_decode(bits, int(v)) for v in np.asarray(q, np.uint64)],
                    dtype=np.float64)


def gather_f32(idx, deg, src, vals) -> np.ndarray:
    idx, deg, src = _u32(idx), _u32(deg), _u32(src)
    vals = np.ascontiguousarray(vals, np.float32)
    out = np.zeros(idx.size, np.float32)
    lib().or_gather_f32(idx.size, _p(idx, u32), _p(deg, u32), _p(src, u32), _p(vals, f32),
                        _p(out, f32))
    return out


def gather_u64(idx, deg, src, vals) -> np.ndarray:
    idx, deg, src = _u32(idx), _u32(deg), _u32(src)
    vals = np.ascontiguousarray(vals, np.uint64)
    out = np.zeros(idx.size, np.uint64)
    lib().or_gather_u64(idx.size, _p(idx, u32), _p(deg, u32), _p(src, u32), _p(vals, u64),
                        _p(out, u64))
    return out


def spmv_f32(idx, deg, src, w, x, threads: int = 0) -> np.ndarray:
    idx, deg, src = _u32(idx), _u32(deg), _u32(src)
    w = None if w is None else np.ascontiguousarray(w, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros(idx.size, np.float32)
    lib().or_spmv_f32(idx.size, _p(idx, u32), _p(deg, u32), _p(src, u32), _p(w, f32),
                      _p(x, f32), _p(y, f32), threads)
    return y


def spmv_fixed(idx, deg, src, wq, wconst, xq, threads: int = 0) -> np.ndarray:
    idx, deg, src = _u32(idx), _u32(deg), _u32(src)
    wq = None if wq is None else _u32(wq)
    xq = _u32(xq)
    y = np.zeros(idx.size, np.uint64)
    lib().or_spmv_fixed(idx.size, _p(idx, u32), _p(deg, u32), _p(src, u32), _p(wq, u32),
                        wconst, _p(xq, u32), _p(y, u64), threads)
    return y


def spmv_push_f32(fwd_idx, fwd_deg, fwd_dst, fw, x, reps: int = 1, threads: int = 0):
    """SpMV push (SPMV.c:752-848): reps products accumulated onto y = 0."""
    fwd_idx, fwd_deg, fwd_dst = _u32(fwd_idx), _u32(fwd_deg), _u32(fwd_dst)
    fw = None if fw is None else np.ascontiguousarray(fw, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros(max(x.size, 1), np.float32)
    lib().or_spmv_push_f32(x.size, _p(fwd_idx, u32), _p(fwd_deg, u32), _p(fwd_dst, u32),
                           _p(fw, f32), _p(x, f32), reps, threads, _p(y, f32))
    return y[:x.size]


def spmv_push_fixed(fwd_idx, fwd_deg, fwd_dst, fw, xq, reps: int = 1, threads: int = 0):
    """Fixed-point SpMV push (SPMV.c:1017-1136): Q32.32, MULFixed64V1 products."""
    fwd_idx, fwd_deg, fwd_dst = _u32(fwd_idx), _u32(fwd_deg), _u32(fwd_dst)
    fw = None if fw is None else np.ascontiguousarray(fw, np.float32)
    xq = np.ascontiguousarray(xq, np.uint64)
    y = np.zeros(max(xq.size, 1), np.uint64)
    lib().or_spmv_push_fixed(xq.size, _p(fwd_idx, u32), _p(fwd_deg, u32), _p(fwd_dst, u32),
                             _p(fw, f32), _p(xq, u64), reps, threads, _p(y, u64))
    return y[:xq.size]


def mul_fixed64(a: int, b: int) -> int:
    return int(lib().or_mul_fixed64_check(a, b))


@dataclass
class PrOut:
    ranks: np.ndarray
    scaled: np.ndarray
    iterations: int
    executed: int
    active: int
    error_total: float
    rank_sum: float
    hist_active: np.ndarray
    hist_err: np.ndarray


def pagerank_pull(outdeg, idx, deg, src, kind: int = 0, max_iterations: int = 20,
                  epsilon: float = 1e-4, stop_rule: int = 0, l1_tolerance: float = 1e-6,
                  init: int = 0, init_ranks=None, threads: int = 0) -> PrOut:
    outdeg, idx, deg, src = _u32(outdeg), _u32(idx), _u32(deg), _u32(src)
    n = idx.size
    ir = None if init_ranks is None else np.ascontiguousarray(init_ranks, np.float32)
    ranks = np.zeros(n, np.float32)
    scaled = np.zeros(n, np.float32)
    ha = np.zeros(max(max_iterations, 1), np.uint32)
    he = np.zeros(max(max_iterations, 1), np.float64)
    o = _PrOut()
    st = lib().or_pagerank_pull(n, _p(outdeg, u32), _p(idx, u32), _p(deg, u32), _p(src, u32),
                                kind, max_iterations, epsilon, stop_rule, l1_tolerance, init,
                                _p(ir, f32), threads, _p(ranks, f32), _p(scaled, f32),
                                _p(ha, u32), _p(he, f64), C.byref(o))
    if st != 0:
        raise RuntimeError(f"or_pagerank_pull failed ({st})")
    return PrOut(ranks, scaled, o.iterations, o.executed, o.active, o.error_total, o.rank_sum,
                 ha[:o.executed], he[:o.executed])


def pagerank_push(outdeg, fwd_idx, fwd_deg, fwd_dst, kind: int = 0, max_iterations: int = 20,
                  epsilon: float = 1e-4, stop_rule: int = 0, l1_tolerance: float = 1e-6,
                  init: int = 0, init_ranks=None, threads: int = 0) -> PrOut:
    """Push directions (pageRank.c:1111-1272, 1992-2117) over the forward CSR."""
    outdeg, fwd_idx, fwd_deg, fwd_dst = _u32(outdeg), _u32(fwd_idx), _u32(fwd_deg), _u32(fwd_dst)
    n = outdeg.size
    ir = None if init_ranks is None else np.ascontiguousarray(init_ranks, np.float32)
    ranks = np.zeros(n, np.float32)
    scaled = np.zeros(n, np.float32)
    ha = np.zeros(max(max_iterations, 1), np.uint32)
    he = np.zeros(max(max_iterations, 1), np.float64)
    o = _PrOut()
    st = lib().or_pagerank_push(n, _p(outdeg, u32), _p(fwd_idx, u32), _p(fwd_deg, u32),
                                _p(fwd_dst, u32), kind, max_iterations, epsilon, stop_rule,
                                l1_tolerance, init, _p(ir, f32), threads, _p(ranks, f32),
                                _p(scaled, f32), _p(ha, u32), _p(he, f64), C.byref(o))
    if st != 0:
        raise RuntimeError(f"or_pagerank_push failed ({st})")
    return PrOut(ranks, scaled, o.iterations, o.executed, o.active, o.error_total, o.rank_sum,
                 ha[:o.executed], he[:o.executed])


def pagerank_dd_pull(outdeg, idx, deg, src, fwd_idx, fwd_deg, fwd_dst,
                     max_iterations: int = 20, epsilon: float = 1e-4, init: int = 0,
                     init_ranks=None, threads: int = 0) -> PrOut:
    """Data-driven (worklist) pull, pageRank.c:2810-2969; forward CSR for the worklist."""
    outdeg, idx, deg, src = _u32(outdeg), _u32(idx), _u32(deg), _u32(src)
    fwd_idx, fwd_deg, fwd_dst = _u32(fwd_idx), _u32(fwd_deg), _u32(fwd_dst)
    n = idx.size
    ir = None if init_ranks is None else np.ascontiguousarray(init_ranks, np.float32)
    ranks = np.zeros(n, np.float32)
    scaled = np.zeros(n, np.float32)
    ha = np.zeros(max(max_iterations, 1), np.uint32)
    he = np.zeros(max(max_iterations, 1), np.float64)
    o = _PrOut()
    st = lib().or_pagerank_dd_pull(n, _p(outdeg, u32), _p(idx, u32), _p(deg, u32),
                                   _p(src, u32), _p(fwd_idx, u32), _p(fwd_deg, u32),
                                   _p(fwd_dst, u32), max_iterations, epsilon, init,
                                   _p(ir, f32), threads, _p(ranks, f32), _p(scaled, f32), _p(ha, u32),
                                   _p(he, f64), C.byref(o))
    if st != 0:
        raise RuntimeError(f"or_pagerank_dd_pull failed ({st})")
    return PrOut(ranks, scaled, o.iterations, o.executed, o.active, o.error_total, o.rank_sum,
                 ha[:o.executed], he[:o.executed])


# kind codes of the quantised pulls (pageRank.c:2119-2676): 100 + code width
KIND_QUANT32, KIND_QUANT16, KIND_QUANT8 = 132, 116, 108


def quantize(c, bits: int):
    """One iteration's quantisation of f32 contributions: (scale, u64 codes)."""
    c = np.ascontiguousarray(c, np.float32)
    codes = np.zeros(max(c.size, 1), np.uint64)
    scale = lib().or_quantize(c.size, _p(c, f32), bits, _p(codes, u64))
    return np.float32(scale), codes[:c.size]


def pagerank_f64(outdeg, idx, deg, src, tol: float = 1e-14, max_iterations: int = 2000,
                 threads: int = 0) -> np.ndarray:
    outdeg, idx, deg, src = _u32(outdeg), _u32(idx), _u32(deg), _u32(src)
    out = np.zeros(idx.size, np.float64)
    lib().or_pagerank_f64(idx.size, _p(outdeg, u32), _p(idx, u32), _p(deg, u32), _p(src, u32),
                          tol, max_iterations, threads, _p(out, f64))
    return out


def ppr_f64(outdeg, idx, deg, src, seed: int, tol: float = 1e-14,
            max_iterations: int = 3000, threads: int = 0) -> np.ndarray:
    """FP64 personalised PageRank column (seed = 0xFFFFFFFF: uniform restart)."""
    outdeg, idx, deg, src = _u32(outdeg), _u32(idx), _u32(deg), _u32(src)
    out = np.zeros(idx.size, np.float64)
    lib().or_ppr_f64(idx.size, _p(outdeg, u32), _p(idx, u32), _p(deg, u32), _p(src, u32),
                     seed, tol, max_iterations, threads, _p(out, f64))
    return out


def ppr_residual(outdeg, idx, deg, src, seeds, ranks, threads: int = 0):
    """(residual L1, column L1) per column of normalised ranks [n, k]: the FP64 residual
    p + 0.85 A r - r of the (personalised) PageRank equation. ||r - r*||_1 <= residual / 0.15."""
    outdeg, idx, deg, src = _u32(outdeg), _u32(idx), _u32(deg), _u32(src)
    seeds = _u32(np.atleast_1d(seeds))
    r = np.ascontiguousarray(ranks, np.float32)
    k = seeds.size
    assert r.size == idx.size * k
    res = np.zeros(k, np.float64)
    nrm = np.zeros(k, np.float64)
    st = lib().or_ppr_residual(idx.size, _p(outdeg, u32), _p(idx, u32), _p(deg, u32),
                               _p(src, u32), k, _p(seeds, u32), _p(r, f32), threads,
                               _p(res, f64), _p(nrm, f64))
    if st != 0:
        raise RuntimeError(f"or_ppr_residual failed ({st})")
    return res, nrm


def build_csr(n: int, src, dst):
    src, dst = _u32(src), _u32(dst)
    m = src.size
    outdeg = np.zeros(n, np.uint32)
    eidx = np.zeros(n, np.uint32)
    fdest = np.zeros(max(m, 1), np.uint32)
    indeg = np.zeros(n, np.uint32)
    iidx = np.zeros(n, np.uint32)
    isrc = np.zeros(max(m, 1), np.uint32)
    mo = C.c_uint64(0)
    st = lib().or_build_csr(n, m, _p(src, u32), _p(dst, u32), _p(outdeg, u32), _p(eidx, u32),
                            _p(fdest, u32), _p(indeg, u32), _p(iidx, u32), _p(isrc, u32),
                            C.byref(mo))
    if st != 0:
        raise RuntimeError(f"or_build_csr failed ({st})")
    k = mo.value
    return dict(out_degree=outdeg, edges_idx=eidx, fwd_dest=fdest[:k], in_degree=indeg,
                inv_idx=iidx, inv_src=isrc[:k], m=k)


def max_threads() -> int:
    return lib().or_max_threads()
