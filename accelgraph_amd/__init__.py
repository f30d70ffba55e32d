"""accelgraph_amd -- MI355X (gfx950) PageRank-pull / SpMV-pull engine.

Python view of the C-ABI in include/accelgraph.h (and of the host graph builder in
include/graphCSR.h). The compute path is libaccelgraph.so: hand-written HIP kernels plus a
C host layer. This module only marshals numpy arrays; it has no compute fallback, and it
raises if the shared library is not built.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from ._abi import (BIN_NAMES, GATHER_F32, GATHER_U8, GATHER_U16, GATHER_U32, GATHER_U64,
                   INIT_REFERENCE, INIT_UNIFORM, INIT_USER, KIND_F32, KIND_FIXED8,
                   KIND_FIXED16, KIND_FIXED32, KIND_FIXED64, KIND_QUANT32, KIND_QUANT16,
                   KIND_QUANT8, DIR_PULL, DIR_PUSH, DIR_DATA_DRIVEN, STOP_ACTIVE, STOP_L1, STOP_NONE,
                   STORE_BF16, STORE_F16, STORE_F32, PPR_UNIFORM, SRC_BLOCK_AUTO, ORDER_FAST, ORDER_REFERENCE,
                   TILES_AUTO, TILES_ROUND_ROBIN, TILES_TICKET, TOP_AUTO, TOP_OFF, TOP_ON,
                   SORT_AUTO, SORT_OFF, SORT_ON,
                   SRC_BLOCK_OFF, FILE_TEXT_EDGE_LIST, FILE_BINARY_EDGE_LIST,
                   FILE_CSR_SERIALIZED, GraphInfo, PlanParams, PrOptions, PrResult)

__all__ = ["lib", "AGError", "Graph", "Device", "DeviceGraph", "PrRun", "Batch", "MLP",
           "plan_params",
           "RMAT_A", "RMAT_B", "RMAT_C", "DEFAULT_SEED"]

RMAT_A, RMAT_B, RMAT_C = 0.57, 0.19, 0.19
DEFAULT_SEED = 27491095  # the reference CLI's MT19937 seed (accel-graph.c:329)

_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = _abi.load()
    return _LIB


class AGError(RuntimeError):
    def __init__(self, status: int, what: str, msg: str):
        super().__init__(f"{what}: {_abi.STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def _check(status: int, what: str) -> None:
    if status != _abi.AG_OK:
        raise AGError(status, what, lib().ag_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray | None, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def _view(ptr, n: int, dtype) -> np.ndarray:
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype)


# ------------------------------------------------------------------ host graph

class Graph:
    """A struct GraphCSR built and owned by the C library (released on close/GC)."""

    def __init__(self, ptr):
        if not ptr:
            raise AGError(_abi.AG_ERR_INVALID, "graph construction", "returned NULL")
        self._p = ptr

    # construction ------------------------------------------------------
    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, a: float = RMAT_A, b: float = RMAT_B,
             c: float = RMAT_C, seed: int = DEFAULT_SEED, simplify: bool = True) -> "Graph":
        return cls(lib().graphCSRGenerateRMAT(scale, edge_factor, a, b, c, seed, int(simplify)))

    @classmethod
    def erdos_renyi(cls, n: int, m: int, seed: int = DEFAULT_SEED,
                    simplify: bool = True) -> "Graph":
        return cls(lib().graphCSRGenerateER(n, m, seed, int(simplify)))

    @classmethod
    def from_edges(cls, src, dst, n: int = 0, weights=None, simplify: bool = True) -> "Graph":
        s = np.ascontiguousarray(src, dtype=np.uint32)
        d = np.ascontiguousarray(dst, dtype=np.uint32)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        if s.shape != d.shape:
            raise ValueError("src and dst lengths differ")
        return cls(lib().graphCSRFromEdgeList(n, s.size, _ptr(s, C.c_uint32),
                                              _ptr(d, C.c_uint32), _ptr(w, C.c_float),
                                              int(simplify)))

    # construction on the GPU: the same arrays as the host builders above ----
    @classmethod
    def rmat_gpu(cls, dev: "Device", scale: int, edge_factor: int = 16, a: float = RMAT_A,
                 b: float = RMAT_B, c: float = RMAT_C, seed: int = DEFAULT_SEED,
                 simplify: bool = True) -> "Graph":
        out = C.POINTER(_abi.GraphCSR)()
        _check(lib().ag_graph_csr_generate_rmat(dev.handle, scale, edge_factor, a, b, c, seed,
                                                int(simplify), C.byref(out)),
               "ag_graph_csr_generate_rmat")
        return cls(out)

    @classmethod
    def erdos_renyi_gpu(cls, dev: "Device", n: int, m: int, seed: int = DEFAULT_SEED,
                        simplify: bool = True) -> "Graph":
        out = C.POINTER(_abi.GraphCSR)()
        _check(lib().ag_graph_csr_generate_er(dev.handle, n, m, seed, int(simplify),
                                              C.byref(out)), "ag_graph_csr_generate_er")
        return cls(out)

    @classmethod
    def from_edges_gpu(cls, dev: "Device", src, dst, n: int = 0, weights=None,
                       simplify: bool = True) -> "Graph":
        s = np.ascontiguousarray(src, dtype=np.uint32)
        d = np.ascontiguousarray(dst, dtype=np.uint32)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        if s.shape != d.shape:
            raise ValueError("src and dst lengths differ")
        out = C.POINTER(_abi.GraphCSR)()
        _check(lib().ag_graph_csr_build(dev.handle, n, s.size, _ptr(s, C.c_uint32),
                                        _ptr(d, C.c_uint32), _ptr(w, C.c_float), int(simplify),
                                        C.byref(out)), "ag_graph_csr_build")
        return cls(out)

    @classmethod
    def load_binary(cls, path: str, simplify: bool = True) -> "Graph":
        return cls(lib().graphCSRLoadBinaryEdgeList(path.encode(), int(simplify)))

    # graph files ---------------------------------------------------------
    @classmethod
    def _loaded(cls, ptr, what: str) -> "Graph":
        if not ptr:
            raise AGError(_abi.AG_ERR_INVALID, what, lib().graphIOLastError().decode())
        return cls(ptr)

    @classmethod
    def load_text(cls, path: str, weighted: bool = False, simplify: bool = True) -> "Graph":
        """Text edge list: `src dst [weight]` per line (space, tab or comma separated)."""
        return cls._loaded(lib().graphCSRLoadTextEdgeList(os.fsencode(path), int(weighted),
                                                          int(simplify)),
                           "graphCSRLoadTextEdgeList")

    @classmethod
    def load_serialized(cls, path: str) -> "Graph":
        return cls._loaded(lib().graphCSRLoadSerialized(os.fsencode(path)),
                           "graphCSRLoadSerialized")

    @classmethod
    def load_file(cls, path: str, file_format: int = FILE_BINARY_EDGE_LIST,
                  simplify: bool = True) -> "Graph":
        """`file_format`: 0 text edge list, 1 binary edge list, 2 serialised CSR."""
        return cls._loaded(lib().graphCSRLoadFile(os.fsencode(path), int(file_format),
                                                  int(simplify)), "graphCSRLoadFile")

    def save(self, path: str, file_format: int = FILE_CSR_SERIALIZED) -> None:
        if lib().graphCSRSaveFile(self.ptr, os.fsencode(path), int(file_format)) != 0:
            raise AGError(_abi.AG_ERR_INVALID, "graphCSRSaveFile",
                          lib().graphIOLastError().decode())

    def set_weights(self, forward_weights) -> None:
        w = np.ascontiguousarray(forward_weights, dtype=np.float32)
        if w.size != self.m:
            raise ValueError("need one weight per forward edge")
        if lib().graphCSRSetWeights(self._p, _ptr(w, C.c_float)) != 0:
            raise AGError(_abi.AG_ERR_OOM, "graphCSRSetWeights", "failed")

    def close(self) -> None:
        if self._p:
            lib().graphCSRFree(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # views (valid while the graph lives) ---------------------------------
    @property
    def ptr(self):
        return self._p

    @property
    def n(self) -> int:
        return self._p.contents.num_vertices

    @property
    def m(self) -> int:
        return self._p.contents.num_edges

    def _v(self, inverse: bool):
        g = self._p.contents
        return (g.inverse_vertices if inverse else g.vertices).contents

    def _e(self, inverse: bool):
        g = self._p.contents
        return (g.inverse_sorted_edges_array if inverse else g.sorted_edges_array).contents

    @property
    def out_degree(self) -> np.ndarray:
        return _view(self._v(False).out_degree, self.n, np.uint32)

    @property
    def edges_idx(self) -> np.ndarray:
        return _view(self._v(False).edges_idx, self.n, np.uint32)

    @property
    def fwd_dest(self) -> np.ndarray:
        return _view(self._e(False).edges_array_dest, self.m, np.uint32)

    @property
    def fwd_src(self) -> np.ndarray:
        return _view(self._e(False).edges_array_src, self.m, np.uint32)

    @property
    def in_degree(self) -> np.ndarray:
        return _view(self._v(True).out_degree, self.n, np.uint32)

    @property
    def inv_idx(self) -> np.ndarray:
        return _view(self._v(True).edges_idx, self.n, np.uint32)

    @property
    def inv_src(self) -> np.ndarray:
        return _view(self._e(True).edges_array_dest, self.m, np.uint32)

    @property
    def inv_weight(self) -> np.ndarray | None:
        p = self._e(True).edges_array_weight
        return _view(p, self.m, np.float32) if p else None

    @property
    def fwd_weight(self) -> np.ndarray | None:
        p = self._e(False).edges_array_weight
        return _view(p, self.m, np.float32) if p else None


def plan_params(**kw) -> PlanParams:
    p = PlanParams()
    lib().ag_plan_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


# ------------------------------------------------------------------ device

class Device:
    def __init__(self, ordinal: int = 0):
        h = C.c_void_p()
        _check(lib().ag_device_open(ordinal, C.byref(h)), "ag_device_open")
        self._h = h
        self._graphs = weakref.WeakSet()

    @staticmethod
    def count() -> int:
        c = C.c_int(0)
        st = lib().ag_device_count(C.byref(c))
        return c.value if st == _abi.AG_OK else 0

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        """Free every graph still uploaded on this device, then close it."""
        if self._h:
            for g in list(self._graphs):
                g.close()
            _check(lib().ag_device_close(self._h), "ag_device_close")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self) -> None:
        _check(lib().ag_device_sync(self._h), "ag_device_sync")

    def copy_bandwidth(self, nbytes: int = 1 << 30, reps: int = 10) -> float:
        """Measured HBM copy bandwidth in GB/s (read + write bytes)."""
        g = C.c_double(0.0)
        _check(lib().ag_device_copy_bandwidth(self._h, nbytes, reps, C.byref(g)),
               "ag_device_copy_bandwidth")
        return g.value

    def set_watchdog(self, run_timeout_ms: int, poll_interval_us: int = 50) -> None:
        _check(lib().ag_device_set_watchdog(self._h, run_timeout_ms, poll_interval_us),
               "ag_device_set_watchdog")

    def upload(self, graph: "Graph | None" = None, *, n=None, out_degree=None, inv_idx=None,
               in_degree=None, inv_src=None, inv_weight=None, weights: bool = False,
               params: PlanParams | None = None) -> "DeviceGraph":
        """Upload a Graph (or raw reference-layout arrays) into HBM."""
        if graph is not None:
            n = graph.n
            out_degree, inv_idx, in_degree, inv_src = (graph.out_degree, graph.inv_idx,
                                                       graph.in_degree, graph.inv_src)
            inv_weight = graph.inv_weight if weights else None
        arrs = [np.ascontiguousarray(a, dtype=np.uint32)
                for a in (out_degree, inv_idx, in_degree, inv_src)]
        w = None if inv_weight is None else np.ascontiguousarray(inv_weight, dtype=np.float32)
        h = C.c_void_p()
        _check(lib().ag_graph_upload(self._h, int(n), int(arrs[3].size),
                                     *[_ptr(a, C.c_uint32) for a in arrs], _ptr(w, C.c_float),
                                     C.byref(params) if params is not None else None,
                                     C.byref(h)), "ag_graph_upload")
        dg = DeviceGraph(self, h, int(n))
        self._graphs.add(dg)
        return dg

    def fixed_encode(self, bits: int, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros(x.size, dtype=np.uint64)
        _check(lib().ag_fixed_encode(self._h, bits, _ptr(x, C.c_float), _ptr(out, C.c_uint64),
                                     x.size), "ag_fixed_encode")
        return out

    def fixed_decode(self, bits: int, q) -> np.ndarray:
        q = np.ascontiguousarray(q, dtype=np.uint64)
        out = np.zeros(q.size, dtype=np.float64)
        _check(lib().ag_fixed_decode(self._h, bits, _ptr(q, C.c_uint64), _ptr(out, C.c_double),
                                     q.size), "ag_fixed_decode")
        return out


@dataclass
class PrRun:
    ranks: np.ndarray
    iterations: int
    executed: int
    active: int
    error_total: float
    rank_sum: float
    loop_ms: float
    total_ms: float
    refine_iteration: int
    hist_active: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    hist_err: np.ndarray = field(default_factory=lambda: np.zeros(0))
    hist_ms: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))


def _options(**kw) -> PrOptions:
    o = PrOptions()
    lib().ag_pr_options_default(C.byref(o))
    for k, v in kw.items():
        if k == "init_ranks":
            continue
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


class DeviceGraph:
    def __init__(self, dev: Device, handle, n: int):
        self.dev = dev
        self._h = handle
        self.n = n

    def close(self) -> None:
        if self._h:
            lib().ag_graph_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def info(self) -> GraphInfo:
        i = GraphInfo()
        _check(lib().ag_graph_get_info(self._h, C.byref(i)), "ag_graph_get_info")
        return i

    def features(self) -> np.ndarray:
        """The warm-start features of every vertex (original order), shape [n, 5]."""
        X = np.zeros((self.n, 5), np.float32)
        _check(lib().ag_mlp_features(self._h, X.ctypes.data_as(C.POINTER(C.c_float))),
               "ag_mlp_features")
        return X

    def pagerank(self, history: bool = True, scaled: bool = False, warm=None, **kw) -> PrRun:
        """Run PageRank pull; kw are ag_pr_options fields (kind, max_iterations, epsilon,
        stop_rule, l1_tolerance, init, init_ranks, storage, refine_l1, chunk). warm = an MLP
        starts the run from its predicted ranks (uniform init, FP32), with the fallback rule."""
        init_ranks = kw.get("init_ranks")
        ir = None
        if init_ranks is not None:
            ir = np.ascontiguousarray(init_ranks, dtype=np.float32)
            kw["init"] = INIT_USER
        o = _options(record_history=int(history), **kw)
        if ir is not None:
            o.init_ranks = _ptr(ir, C.c_float)
        s = C.c_void_p()
        _check(lib().ag_pr_begin(self._h, C.byref(o), C.byref(s)), "ag_pr_begin")
        r = PrResult()
        # iterations queued per poll: ag_pagerank_pull's automatic rule (32 up to 2^24 edges,
        # where an iteration costs about what a poll does; else 8)
        maxit = o.max_iterations
        chunk = o.chunk or (32 if self.info().num_edges <= (1 << 24) else 8)
        enq = 0
        wres = None
        if warm is not None:
            wres = _abi.WarmResult()
            st = lib().ag_pr_warm_start(s, warm.handle, C.byref(wres))
            if st != _abi.AG_OK:
                lib().ag_pr_end(s, None, None)
                _check(st, "ag_pr_warm_start")
            enq = 3
        try:
            while enq < maxit:
                k = min(chunk, maxit - enq)
                _check(lib().ag_pr_step(s, k), "ag_pr_step")
                _check(lib().ag_pr_poll(s, C.byref(r)), "ag_pr_poll")
                enq += k
                if r.executed > r.iterations:
                    break
            ha = np.zeros(max(r.executed, 1), np.uint32)
            he = np.zeros(max(r.executed, 1), np.float64)
            hm = np.zeros(max(r.executed, 1), np.float32)
            rows = C.c_uint32(0)
            if history:
                _check(lib().ag_pr_history(s, r.executed, _ptr(ha, C.c_uint32),
                                           _ptr(he, C.c_double), _ptr(hm, C.c_float),
                                           C.byref(rows)), "ag_pr_history")
            sc = np.zeros(self.n, np.float32) if scaled else None
            if scaled:
                _check(lib().ag_pr_read_scaled(s, _ptr(sc, C.c_float)), "ag_pr_read_scaled")
        except Exception:
            lib().ag_pr_end(s, None, None)
            raise
        ranks = np.zeros(self.n, np.float32)
        _check(lib().ag_pr_end(s, _ptr(ranks, C.c_float), C.byref(r)), "ag_pr_end")
        run = PrRun(ranks=ranks, iterations=r.iterations, executed=r.executed,
                    active=r.active_vertices, error_total=r.error_total, rank_sum=r.rank_sum,
                    loop_ms=r.loop_ms, total_ms=r.total_ms, refine_iteration=r.refine_iteration,
                    hist_active=ha[:rows.value], hist_err=he[:rows.value],
                    hist_ms=hm[:rows.value])
        if scaled:
            run.scaled = sc
        run.warm = wres
        return run

    def ppr(self, seeds, **kw):
        """Personalised PageRank, one column per seed (PPR_UNIFORM = plain restart).
        Returns (ranks [n, k], column L1 deltas, PprResult)."""
        sd = np.ascontiguousarray(seeds, dtype=np.uint32)
        k = sd.size
        o = _abi.PprOptions()
        lib().ag_ppr_options_default(C.byref(o))
        for key, v in kw.items():
            setattr(o, key, v)
        out = np.zeros(self.n * k, np.float32)
        ce = np.zeros(k, np.float64)
        r = _abi.PprResult()
        _check(lib().ag_ppr_pull(self._h, k, _ptr(sd, C.c_uint32), C.byref(o),
                                 _ptr(out, C.c_float), _ptr(ce, C.c_double), C.byref(r)),
               "ag_ppr_pull")
        return out.reshape(self.n, k), ce, r

    def session(self, **kw) -> "PrSession":
        return PrSession(self, **kw)

    def spmv(self, x, reps: int = 1) -> tuple[np.ndarray, float]:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(self.n, np.float32)
        ms = C.c_double(0.0)
        _check(lib().ag_spmv_pull(self._h, _ptr(x, C.c_float), _ptr(y, C.c_float), reps,
                                  C.byref(ms)), "ag_spmv_pull")
        return y, ms.value

    def spmv_fixed(self, xq, reps: int = 1) -> tuple[np.ndarray, float]:
        xq = np.ascontiguousarray(xq, dtype=np.uint32)
        y = np.zeros(self.n, np.uint64)
        ms = C.c_double(0.0)
        _check(lib().ag_spmv_pull_fixed(self._h, _ptr(xq, C.c_uint32), _ptr(y, C.c_uint64),
                                        reps, C.byref(ms)), "ag_spmv_pull_fixed")
        return y, ms.value

    def spmv_push(self, x, reps: int = 1) -> tuple[np.ndarray, float]:
        """Push form: `reps` products accumulated onto y = 0 (the reference's push)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(self.n, np.float32)
        ms = C.c_double(0.0)
        _check(lib().ag_spmv_push(self._h, _ptr(x, C.c_float), _ptr(y, C.c_float), reps,
                                  C.byref(ms)), "ag_spmv_push")
        return y, ms.value

    def spmv_push_fixed(self, xq, reps: int = 1) -> tuple[np.ndarray, float]:
        """Fixed-point push: Q32.32 x and y."""
        xq = np.ascontiguousarray(xq, dtype=np.uint64)
        y = np.zeros(self.n, np.uint64)
        ms = C.c_double(0.0)
        _check(lib().ag_spmv_push_fixed(self._h, _ptr(xq, C.c_uint64), _ptr(y, C.c_uint64),
                                        reps, C.byref(ms)), "ag_spmv_push_fixed")
        return y, ms.value

    def gather(self, values, gtype: int) -> np.ndarray:
        vt = {GATHER_F32: np.float32, GATHER_U64: np.uint64, GATHER_U32: np.uint32,
              GATHER_U16: np.uint16, GATHER_U8: np.uint8}[gtype]
        v = np.ascontiguousarray(values, dtype=vt)
        out = np.zeros(self.n, np.float32 if gtype == GATHER_F32 else np.uint64)
        _check(lib().ag_pull_gather(self._h, gtype, v.ctypes.data_as(C.c_void_p),
                                    out.ctypes.data_as(C.c_void_p)), "ag_pull_gather")
        return out


class Batch:
    """Many independent graphs solved by one persistent (hipGraph-captured) launch."""

    def __init__(self, dev: "Device", graphs):
        self.dev = dev
        G = len(graphs)
        vo = np.zeros(G + 1, np.uint32)
        eo = np.zeros(G + 1, np.uint32)
        for i, g in enumerate(graphs):
            vo[i + 1] = vo[i] + g.n
            eo[i + 1] = eo[i] + g.m
        cat = lambda f: (np.concatenate([np.asarray(f(g), np.uint32) for g in graphs])
                         if vo[-1] else np.zeros(0, np.uint32))
        od, ii, dg = cat(lambda g: g.out_degree), cat(lambda g: g.inv_idx), cat(lambda g: g.in_degree)
        src = (np.concatenate([np.asarray(g.inv_src, np.uint32) for g in graphs])
               if eo[-1] else np.zeros(0, np.uint32))
        self.vert_off, self.edge_off = vo, eo
        h = C.c_void_p()
        _check(lib().ag_batch_upload(dev.handle, G, _ptr(vo, C.c_uint32), _ptr(eo, C.c_uint32),
                                     _ptr(od, C.c_uint32), _ptr(ii, C.c_uint32),
                                     _ptr(dg, C.c_uint32), _ptr(src, C.c_uint32), C.byref(h)),
               "ag_batch_upload")
        self._h = h
        dev._graphs.add(self)

    def close(self) -> None:
        if self._h:
            lib().ag_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pagerank(self, **kw):
        o = _abi.BatchOptions()
        lib().ag_batch_options_default(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        G = self.vert_off.size - 1
        ranks = np.zeros(int(self.vert_off[-1]), np.float32)
        its = np.zeros(G, np.uint32)
        err = np.zeros(G, np.float64)
        ms = C.c_double(0.0)
        _check(lib().ag_batch_pagerank(self._h, C.byref(o), _ptr(ranks, C.c_float),
                                       _ptr(its, C.c_uint32), _ptr(err, C.c_double),
                                       C.byref(ms)), "ag_batch_pagerank")
        per = [ranks[self.vert_off[i]:self.vert_off[i + 1]] for i in range(G)]
        return per, its, err, ms.value


class MLP:
    """The warm-start network on the device (see include/accelgraph.h, learned warm start)."""

    def __init__(self, dev: "Device", seed: int = 0):
        self.dev = dev
        h = C.c_void_p()
        _check(lib().ag_mlp_create(dev.handle, seed, C.byref(h)), "ag_mlp_create")
        self._h = h
        self.P = lib().ag_mlp_param_count()
        self.norm = (np.zeros(5, np.float32), np.ones(5, np.float32), 0.0, 1.0)
        dev._graphs.add(self)

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            lib().ag_mlp_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.P, np.float32)
        _check(lib().ag_mlp_get_params(self._h, _ptr(p, C.c_float)), "ag_mlp_get_params")
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.P
        _check(lib().ag_mlp_set_params(self._h, _ptr(p, C.c_float)), "ag_mlp_set_params")

    def set_normalization(self, in_mean, in_std, y_mean, y_std) -> None:
        m = np.ascontiguousarray(in_mean, np.float32)
        s = np.ascontiguousarray(in_std, np.float32)
        _check(lib().ag_mlp_set_normalization(self._h, _ptr(m, C.c_float), _ptr(s, C.c_float),
                                              float(y_mean), float(y_std)),
               "ag_mlp_set_normalization")
        self.norm = (m.copy(), s.copy(), float(y_mean), float(y_std))

    def fit_normalization(self, X, y) -> None:
        X = np.asarray(X, np.float64)
        y = np.asarray(y, np.float64)
        self.set_normalization(X.mean(0), np.maximum(X.std(0), 1e-3), y.mean(),
                               max(y.std(), 1e-3))

    def loss_grad(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros(self.P, np.float32)
        loss = C.c_double(0.0)
        _check(lib().ag_mlp_loss_grad(self._h, _ptr(X, C.c_float), _ptr(y, C.c_float),
                                      y.size, C.byref(loss), _ptr(g, C.c_float)),
               "ag_mlp_loss_grad")
        return loss.value, g

    def train(self, X, y, epochs: int = 5, batch: int = 8192, lr: float = 2e-3,
              seed: int = 0) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        losses = np.zeros(epochs, np.float32)
        _check(lib().ag_mlp_train(self._h, _ptr(X, C.c_float), _ptr(y, C.c_float), y.size,
                                  epochs, batch, lr, seed, _ptr(losses, C.c_float)),
               "ag_mlp_train")
        return losses

    def predict(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros(X.shape[0], np.float32)
        _check(lib().ag_mlp_predict(self._h, _ptr(X, C.c_float), X.shape[0],
                                    _ptr(out, C.c_float)), "ag_mlp_predict")
        return out

    def save(self, path: str) -> None:
        import json
        m, s, ym, ys = self.norm
        with open(path, "w") as f:
            json.dump({"params": [float(v) for v in self.get_params()],
                       "in_mean": [float(v) for v in m], "in_std": [float(v) for v in s],
                       "y_mean": ym, "y_std": ys}, f)

    def load(self, path: str) -> None:
        import json
        with open(path) as f:
            d = json.load(f)
        self.set_params(np.array(d["params"], np.float32))
        self.set_normalization(d["in_mean"], d["in_std"], d["y_mean"], d["y_std"])


class PrSession:
    """Stepping interface (ag_pr_begin / ag_pr_step / ag_pr_poll / ag_pr_end)."""

    def __init__(self, dg: DeviceGraph, **kw):
        self.dg = dg
        ir = kw.pop("init_ranks", None)
        self._ir = None
        if ir is not None:
            self._ir = np.ascontiguousarray(ir, dtype=np.float32)
            kw["init"] = INIT_USER
        self.opts = _options(**kw)
        if self._ir is not None:
            self.opts.init_ranks = _ptr(self._ir, C.c_float)
        s = C.c_void_p()
        _check(lib().ag_pr_begin(dg.handle, C.byref(self.opts), C.byref(s)), "ag_pr_begin")
        self._s = s

    def step(self, k: int) -> None:
        _check(lib().ag_pr_step(self._s, k), "ag_pr_step")

    def poll(self) -> PrResult:
        r = PrResult()
        _check(lib().ag_pr_poll(self._s, C.byref(r)), "ag_pr_poll")
        return r

    def last_step_kernel_ms(self) -> tuple[float, int]:
        ms, n = C.c_double(0.0), C.c_uint32(0)
        _check(lib().ag_pr_last_step_kernel_ms(self._s, C.byref(ms), C.byref(n)),
               "ag_pr_last_step_kernel_ms")
        return ms.value, n.value

    def end(self) -> tuple[np.ndarray, PrResult]:
        ranks = np.zeros(self.dg.n, np.float32)
        r = PrResult()
        s, self._s = self._s, None
        _check(lib().ag_pr_end(s, _ptr(ranks, C.c_float), C.byref(r)), "ag_pr_end")
        return ranks, r

    def __del__(self):
        if getattr(self, "_s", None):
            try:
                lib().ag_pr_end(self._s, None, None)
            except Exception:
                pass
