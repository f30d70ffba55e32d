"""ctypes declarations of the C-ABI (include/accelgraph.h, include/graphCSR.h,
include/pageRank.h, include/SPMV.h, include/graphConfig.h).

Struct layouts here must match the headers field for field; tests/test_abi.py checks the
sizes against the compiled library's view where it can.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libaccelgraph.so")

u32, u64, i32, f32, f64 = C.c_uint32, C.c_uint64, C.c_int32, C.c_float, C.c_double
P = C.POINTER

AG_OK = 0
STATUS_NAMES = {0: "AG_OK", 1: "AG_ERR_INVALID", 2: "AG_ERR_NO_DEVICE", 3: "AG_ERR_HIP",
                4: "AG_ERR_OOM", 5: "AG_ERR_TIMEOUT", 6: "AG_ERR_UNSUPPORTED"}
AG_ERR_INVALID, AG_ERR_NO_DEVICE, AG_ERR_HIP, AG_ERR_OOM, AG_ERR_TIMEOUT, AG_ERR_UNSUPPORTED = \
    1, 2, 3, 4, 5, 6

AG_NBINS = 5
SRC_BLOCK_AUTO, SRC_BLOCK_OFF = 0, 255
ORDER_FAST, ORDER_REFERENCE = 0, 1
TILES_AUTO, TILES_ROUND_ROBIN, TILES_TICKET = 0, 1, 2
TOP_AUTO, TOP_OFF, TOP_ON = 0, 1, 2
SORT_AUTO, SORT_OFF, SORT_ON = 0, 1, 2
# graphCSR.h GRAPH_FILE_*: the reference driver's -z codes
FILE_TEXT_EDGE_LIST, FILE_BINARY_EDGE_LIST, FILE_CSR_SERIALIZED = 0, 1, 2
BIN_NAMES = ["thread", "group16", "wave", "block", "hub"]
STOP_ACTIVE, STOP_L1, STOP_NONE = 0, 1, 2
INIT_REFERENCE, INIT_UNIFORM, INIT_USER = 0, 1, 2
KIND_F32, KIND_FIXED64, KIND_FIXED32, KIND_FIXED16, KIND_FIXED8 = 0, 64, 32, 16, 8
KIND_QUANT32, KIND_QUANT16, KIND_QUANT8 = 132, 116, 108
DIR_PULL, DIR_PUSH, DIR_DATA_DRIVEN = 0, 1, 2
STORE_F32, STORE_BF16, STORE_F16 = 0, 1, 2
GATHER_F32, GATHER_U64, GATHER_U32, GATHER_U16, GATHER_U8 = 0, 1, 2, 3, 4
PPR_UNIFORM = 0xFFFFFFFF


class PlanParams(C.Structure):
    _fields_ = [("thread_max_deg", u32), ("group16_max_deg", u32), ("wave_max_deg", u32),
                ("block_max_deg", u32), ("tile_rows", u32), ("tile_nnz", u32),
                ("segment_nnz", u32), ("order", u32), ("src_block_bits", u32),
                ("wg_per_cu", u32), ("blocked_min_bin", u32), ("hot_segments", u32),
                ("fp32_order", u32), ("tile_order", u32),
                ("piece_nnz", u32), ("top_table", u32), ("tile_sort", u32)]


class GraphInfo(C.Structure):
    _fields_ = [("num_vertices", u32), ("num_edges", u32), ("bin_rows", u32 * AG_NBINS),
                ("bin_edges", u64 * AG_NBINS), ("bin_blocks", u32 * AG_NBINS),
                ("num_tiles", u32), ("num_segments", u32), ("has_weights", u32),
                ("device_bytes", u64), ("params", PlanParams), ("src_block_bits", u32),
                ("src_blocks", u32), ("virtual_rows", u32), ("blocked_row0", u32),
                ("hot_segments", u32), ("hot_segment_bits", u32), ("hot_edges", u64),
                ("top_sources", u32), ("top_edges", u64), ("sorted_tiles", u32)]


class PrOptions(C.Structure):
    _fields_ = [("kind", u32), ("max_iterations", u32), ("epsilon", f64), ("stop_rule", u32),
                ("l1_tolerance", f64), ("init", u32), ("init_ranks", P(f32)),
                ("storage", u32), ("refine_l1", f64), ("chunk", u32),
                ("record_history", u32), ("direction", u32)]


class PrResult(C.Structure):
    _fields_ = [("iterations", u32), ("active_vertices", u32), ("error_total", f64),
                ("rank_sum", f64), ("loop_ms", f64), ("total_ms", f64), ("executed", u32),
                ("refine_iteration", i32)]


class BatchOptions(C.Structure):
    _fields_ = [("max_iterations", u32), ("stop_rule", u32), ("epsilon", f64),
                ("l1_tolerance", f64), ("init", u32), ("use_hip_graph", u32),
                ("part_edges", u32), ("heavy_deg", u32)]


class PprOptions(C.Structure):
    _fields_ = [("max_iterations", u32), ("stop_rule", u32), ("l1_tolerance", f64),
                ("storage", u32), ("refine_l1", f64), ("chunk", u32), ("use_hip_graph", u32),
                ("wg_per_cu", u32)]


class PprResult(C.Structure):
    _fields_ = [("iterations", u32), ("executed", u32), ("max_error", f64), ("loop_ms", f64),
                ("refine_iteration", i32)]


class WarmResult(C.Structure):
    _fields_ = [("fallback", u32), ("uniform_l1", f64), ("warm_l1", f64), ("ms", f64)]


class Vertex(C.Structure):
    _fields_ = [("out_degree", P(u32)), ("in_degree", P(u32)), ("edges_idx", P(u32))]


class EdgeList(C.Structure):
    _fields_ = [("num_edges", u32), ("num_vertices", u32), ("edges_array_src", P(u32)),
                ("edges_array_dest", P(u32)), ("edges_array_weight", P(f32)),
                ("max_weight", f32)]


class GraphCSR(C.Structure):
    _fields_ = [("num_edges", u32), ("num_vertices", u32), ("max_weight", f32),
                ("avg_degree", u32), ("vertices", P(Vertex)),
                ("sorted_edges_array", P(EdgeList)), ("inverse_vertices", P(Vertex)),
                ("inverse_sorted_edges_array", P(EdgeList))]


class MT19937State(C.Structure):
    _fields_ = [("mt", u32 * 624), ("mti", u32)]


class Arguments(C.Structure):
    _fields_ = [("wflag", C.c_int), ("xflag", C.c_int), ("sflag", C.c_int), ("Sflag", C.c_int),
                ("dflag", C.c_int), ("binSize", u32), ("verbosity", u32), ("iterations", u32),
                ("trials", u32), ("epsilon", f64), ("source", C.c_int), ("algorithm", u32),
                ("datastructure", u32), ("pushpull", u32), ("sort", u32), ("lmode", u32),
                ("lmode_l2", u32), ("lmode_l3", u32), ("mmode", u32), ("symmetric", u32),
                ("weighted", u32), ("delta", u32), ("pre_numThreads", C.c_int),
                ("algo_numThreads", C.c_int), ("ker_numThreads", C.c_int),
                ("fnameb", C.c_char_p), ("fnamel", C.c_char_p), ("fnameb_format", u32),
                ("convert_format", u32), ("mt19937var", MT19937State), ("cache_size", u32),
                ("afu_config", u64), ("afu_config_2", u64), ("cu_config", u64),
                ("cu_config_2", u64), ("cu_config_3", u64), ("cu_config_4", u64)]


class PageRankStats(C.Structure):
    _fields_ = [("damp", f32), ("base_pr", f32), ("iterations", u32), ("num_vertices", u32),
                ("activeVertices", u32), ("realRanks", P(u32)), ("pageRanks", P(f32)),
                ("time_total", f64), ("error_total", f64)]


class SPMVStats(C.Structure):
    _fields_ = [("iterations", u32), ("arguments", P(Arguments)), ("num_vertices", u32),
                ("vector_output", P(f32)), ("vector_input", P(f32)), ("time_total", f64)]


VP = C.c_void_p

# name -> (restype, argtypes); every symbol include/*.h declares
SIGNATURES = {
    # accelgraph.h
    "ag_device_count": (C.c_int, [P(C.c_int)]),
    "ag_device_open": (C.c_int, [C.c_int, P(VP)]),
    "ag_device_close": (C.c_int, [VP]),
    "ag_device_sync": (C.c_int, [VP]),
    "ag_status_string": (C.c_char_p, [C.c_int]),
    "ag_last_error": (C.c_char_p, []),
    "ag_abi_version": (C.c_int, []),
    "ag_device_set_watchdog": (C.c_int, [VP, u64, u32]),
    "ag_device_copy_bandwidth": (C.c_int, [VP, u64, u32, P(f64)]),
    "ag_plan_params_default": (None, [P(PlanParams)]),
    "ag_graph_upload": (C.c_int, [VP, u32, u32, P(u32), P(u32), P(u32), P(u32), P(f32),
                                  P(PlanParams), P(VP)]),
    "ag_graph_free": (C.c_int, [VP]),
    "ag_graph_set_fixed_weights": (C.c_int, [VP, P(u32), P(u32), P(u32)]),
    "ag_graph_get_info": (C.c_int, [VP, P(GraphInfo)]),
    "ag_autotune_plan": (C.c_int, [VP, u32, u32, P(u32), P(u32), P(u32), P(u32),
                                   P(PlanParams), u32, u32, u32, P(f64), P(u32)]),
    "ag_pr_options_default": (None, [P(PrOptions)]),
    "ag_pagerank_pull": (C.c_int, [VP, P(PrOptions), P(f32), P(PrResult)]),
    "ag_pr_begin": (C.c_int, [VP, P(PrOptions), P(VP)]),
    "ag_pr_step": (C.c_int, [VP, u32]),
    "ag_pr_poll": (C.c_int, [VP, P(PrResult)]),
    "ag_pr_end": (C.c_int, [VP, P(f32), P(PrResult)]),
    "ag_pr_history": (C.c_int, [VP, u32, P(u32), P(f64), P(f32), P(u32)]),
    "ag_pr_read_scaled": (C.c_int, [VP, P(f32)]),
    "ag_pr_last_step_kernel_ms": (C.c_int, [VP, P(f64), P(u32)]),
    "ag_spmv_pull": (C.c_int, [VP, P(f32), P(f32), u32, P(f64)]),
    "ag_spmv_pull_fixed": (C.c_int, [VP, P(u32), P(u64), u32, P(f64)]),
    "ag_spmv_pull_noalloc": (C.c_int, [VP, P(f32), P(f32)]),
    "ag_spmv_push": (C.c_int, [VP, P(f32), P(f32), u32, P(f64)]),
    "ag_spmv_push_fixed": (C.c_int, [VP, P(u64), P(u64), u32, P(f64)]),
    "ag_pull_gather": (C.c_int, [VP, u32, VP, VP]),
    "ag_ppr_options_default": (None, [P(PprOptions)]),
    "ag_ppr_pull": (C.c_int, [VP, u32, P(u32), P(PprOptions), P(f32), P(f64), P(PprResult)]),
    "ag_mlp_param_count": (u32, []),
    "ag_mlp_create": (C.c_int, [VP, u32, P(VP)]),
    "ag_mlp_free": (C.c_int, [VP]),
    "ag_mlp_set_params": (C.c_int, [VP, P(f32)]),
    "ag_mlp_get_params": (C.c_int, [VP, P(f32)]),
    "ag_mlp_set_normalization": (C.c_int, [VP, P(f32), P(f32), f32, f32]),
    "ag_mlp_loss_grad": (C.c_int, [VP, P(f32), P(f32), u32, P(f64), P(f32)]),
    "ag_mlp_train": (C.c_int, [VP, P(f32), P(f32), u32, u32, u32, f32, u32, P(f32)]),
    "ag_mlp_predict": (C.c_int, [VP, P(f32), u32, P(f32)]),
    "ag_mlp_features": (C.c_int, [VP, P(f32)]),
    "ag_pr_warm_start": (C.c_int, [VP, VP, P(WarmResult)]),
    "ag_batch_upload": (C.c_int, [VP, u32, P(u32), P(u32), P(u32), P(u32), P(u32), P(u32),
                                  P(VP)]),
    "ag_batch_free": (C.c_int, [VP]),
    "ag_batch_options_default": (None, [P(BatchOptions)]),
    "ag_batch_pagerank": (C.c_int, [VP, P(BatchOptions), P(f32), P(u32), P(f64), P(f64)]),
    "ag_fixed_encode": (C.c_int, [VP, u32, P(f32), P(u64), u32]),
    "ag_fixed_decode": (C.c_int, [VP, u32, P(u64), P(f64), u32]),
    # graphCSR.h
    "graphCSRFromEdgeList": (P(GraphCSR), [u32, u64, P(u32), P(u32), P(f32), C.c_int]),
    "graphCSRGenerateRMAT": (P(GraphCSR), [u32, u32, f64, f64, f64, u32, C.c_int]),
    "graphCSRGenerateER": (P(GraphCSR), [u32, u64, u32, C.c_int]),
    "graphGenerateRMATEdges": (C.c_int, [u32, u64, f64, f64, f64, u32, P(u32), P(u32)]),
    "graphCSRLoadBinaryEdgeList": (P(GraphCSR), [C.c_char_p, C.c_int]),
    "graphCSRSetWeights": (C.c_int, [P(GraphCSR), P(f32)]),
    "graphCSRLoadTextEdgeList": (P(GraphCSR), [C.c_char_p, C.c_int, C.c_int]),
    "graphCSRLoadSerialized": (P(GraphCSR), [C.c_char_p]),
    "graphCSRLoadFile": (P(GraphCSR), [C.c_char_p, C.c_int, C.c_int]),
    "graphCSRSaveTextEdgeList": (C.c_int, [P(GraphCSR), C.c_char_p]),
    "graphCSRSaveBinaryEdgeList": (C.c_int, [P(GraphCSR), C.c_char_p]),
    "graphCSRSaveSerialized": (C.c_int, [P(GraphCSR), C.c_char_p]),
    "graphCSRSaveFile": (C.c_int, [P(GraphCSR), C.c_char_p, C.c_int]),
    "graphIOLastError": (C.c_char_p, []),
    "graphCSRFree": (None, [P(GraphCSR)]),
    "graphCSRAlloc": (P(GraphCSR), [u32, u32, C.c_int]),
    "ag_generate_rmat_edges": (C.c_int, [VP, u32, u64, f64, f64, f64, u32, P(u32), P(u32)]),
    "ag_generate_er_edges": (C.c_int, [VP, u32, u64, u32, P(u32), P(u32)]),
    "ag_graph_csr_build": (C.c_int, [VP, u32, u64, P(u32), P(u32), P(f32), C.c_int,
                                     P(P(GraphCSR))]),
    "ag_graph_csr_generate_rmat": (C.c_int, [VP, u32, u32, f64, f64, f64, u32, C.c_int,
                                             P(P(GraphCSR))]),
    "ag_graph_csr_generate_er": (C.c_int, [VP, u32, u64, u32, C.c_int, P(P(GraphCSR))]),
    # mt19937.h
    "initializeMersenneState": (None, [P(MT19937State), u32]),
    "generateRandInt": (u32, [P(MT19937State)]),
    "generateRandDouble": (f64, [P(MT19937State)]),
    # pageRank.h
    "newPageRankStatsGraphCSR": (P(PageRankStats), [P(GraphCSR)]),
    "freePageRankStats": (None, [P(PageRankStats)]),
    "pageRankGraphCSR": (P(PageRankStats), [P(Arguments), P(GraphCSR)]),
    "pageRankPullGraphCSR": (P(PageRankStats), [P(Arguments), P(GraphCSR)]),
    "pageRankPullFixedPoint64BitGraphCSR": (P(PageRankStats), [P(Arguments), P(GraphCSR)]),
    "pageRankPullFixedPoint32BitGraphCSR": (P(PageRankStats), [P(Arguments), P(GraphCSR)]),
    "pageRankPullFixedPoint16BitGraphCSR": (P(PageRankStats), [P(Arguments), P(GraphCSR)]),
    "pageRankPullFixedPoint8BitGraphCSR": (P(PageRankStats), [P(Arguments), P(GraphCSR)]),
    "pageRankPullQuant32BitGraphCSR": (P(PageRankStats), [P(Arguments), P(GraphCSR)]),
    "pageRankPullQuant16BitGraphCSR": (P(PageRankStats), [P(Arguments), P(GraphCSR)]),
    "pageRankPullQuant8BitGraphCSR": (P(PageRankStats), [P(Arguments), P(GraphCSR)]),
    "pageRankPushGraphCSR": (P(PageRankStats), [P(Arguments), P(GraphCSR)]),
    "SPMVPushGraphCSR": (P(SPMVStats), [P(Arguments), P(GraphCSR)]),
    "SPMVPushFixedPointGraphCSR": (P(SPMVStats), [P(Arguments), P(GraphCSR)]),
    "pageRankPushFixedPointGraphCSR": (P(PageRankStats), [P(Arguments), P(GraphCSR)]),
    "pageRankDataDrivenPullGraphCSR": (P(PageRankStats), [P(Arguments), P(GraphCSR)]),
    # SPMV.h
    "newSPMVStatsGraphCSR": (P(SPMVStats), [P(GraphCSR)]),
    "freeSPMVStats": (None, [P(SPMVStats)]),
    "SPMVGraphCSR": (P(SPMVStats), [P(Arguments), P(GraphCSR)]),
    "SPMVPullGraphCSR": (P(SPMVStats), [P(Arguments), P(GraphCSR)]),
    "SPMVPullFixedPointGraphCSR": (P(SPMVStats), [P(Arguments), P(GraphCSR)]),
    # accelgraph_compat.h
    "accelgraphReleaseGraph": (None, []),
    "accelgraphShutdown": (None, []),
}


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libaccelgraph.so and bind every declared symbol. Raises if the library is
    missing: there is no fallback implementation."""
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built; run `python -m accelgraph_amd.build` (needs hipcc)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
