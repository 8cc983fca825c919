"""ctypes binding of libplato_hip.so (declarations in include/plato_hip.h).

The shared library is built in-tree by ``python -m plato_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PLATO_HIP_LIB") or os.path.join(HERE, "libplato_hip.so")

PE_OK = 0
PE_ERR = {1: "invalid argument", 2: "HIP error", 3: "out of memory", 4: "bad state",
          5: "out of range", 6: "internal error"}
PE_F32, PE_F64, PE_U32 = 0, 1, 2
PE_KT_PULL_TILE, PE_KT_PULL_FIXUP, PE_KT_BFS_PUSH, PE_KT_BFS_PULL, PE_KT_PPR_BLOCK = 0, 1, 2, 3, 4
PE_KT_PULL_BIN_SCATTER, PE_KT_PULL_BIN_GATHER = 5, 6
PE_PULL_MERGE_PATH, PE_PULL_BINNED = 0, 1
PE_GRAPH_SYMMETRIZE, PE_GRAPH_BUILD_CSR, PE_GRAPH_WEIGHTED, PE_GRAPH_RELABEL_DEGREE = 1, 2, 4, 8
PE_GRAPH_ENCODE_IDS = 16
PE_GRAPH_SAME_IDS = 32
PE_SPREAD_COUNT, PE_SPREAD_ADD = 0, 1
PE_NSTEP_IN, PE_NSTEP_OUT, PE_NSTEP_BOTH = 1, 2, 3


class PlatoHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{PE_ERR.get(code, code)}: {msg}")
        self.code = code


class GraphInfo(C.Structure):
    _fields_ = [("V", C.c_uint32), ("max_v_i", C.c_uint32), ("vertices", C.c_uint32),
                ("max_in_degree", C.c_uint32), ("edges", C.c_uint64), ("nnz", C.c_uint64),
                ("is_directed", C.c_int32), ("has_csr", C.c_int32), ("has_weights", C.c_int32),
                ("relabeled", C.c_int32), ("tile_items", C.c_uint32), ("num_tiles", C.c_uint64),
                ("num_fixup_rows", C.c_uint64), ("device_bytes", C.c_uint64),
                ("pull_layout", C.c_int32), ("binned_runs", C.c_uint32), ("binned_bytes", C.c_uint64)]


class TrustRankOpts(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("damping", C.c_double), ("seed_num", C.c_uint32),
                ("select_method", C.c_uint32), ("eps", C.c_double), ("random_seed", C.c_uint64),
                ("precision", C.c_int32)]


class BinnedOpts(C.Structure):
    _fields_ = [("src_bits", C.c_uint32), ("dst_bits", C.c_uint32), ("run_cap", C.c_uint32),
                ("unit_log2", C.c_uint32)]


class ApplyParams(C.Structure):
    _fields_ = [("damping", C.c_double), ("base", C.c_double), ("seed_mass", C.c_double),
                ("seeds", C.c_void_p), ("divide_by_od", C.c_int32), ("final_iter", C.c_int32)]


class RankOpts(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("damping", C.c_double), ("eps", C.c_double),
                ("precision", C.c_int32), ("weighted", C.c_int32), ("l1_norm_tol", C.c_double)]


class HyperAnfOpts(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("bits", C.c_uint32)]


class CcStats(C.Structure):
    _fields_ = [("epochs", C.c_uint32), ("components", C.c_uint32), ("seconds", C.c_double)]


class KcoreStats(C.Structure):
    _fields_ = [("epochs", C.c_uint32), ("modified", C.c_uint64), ("seconds", C.c_double)]


class RankStats(C.Structure):
    _fields_ = [("iterations_run", C.c_uint32), ("last_delta", C.c_double),
                ("seconds", C.c_double), ("gteps", C.c_double)]


class BfsStats(C.Structure):
    _fields_ = [("epochs", C.c_uint32), ("visited", C.c_uint32), ("reached_edges", C.c_uint64),
                ("seconds", C.c_double), ("alg_bytes", C.c_uint64)]


class PprBatchOpts(C.Structure):
    _fields_ = [("k", C.c_uint32), ("damping", C.c_double), ("tol", C.c_double),
                ("max_iters", C.c_uint32), ("task_edges", C.c_uint32),
                ("plan_packed", C.c_uint32), ("plan_edges4", C.c_uint32), ("plan_edges2", C.c_uint32)]


PE_BATCH_PLAN_OFF = 0xFFFFFFFF


class ServerOpts(C.Structure):
    _fields_ = [("damping", C.c_double), ("tol", C.c_double), ("max_iters", C.c_uint32),
                ("topk", C.c_uint32), ("iters_per_pass", C.c_uint32), ("max_batches", C.c_uint32),
                ("hbm_reserve_bytes", C.c_uint64)]


class RequestStats(C.Structure):
    _fields_ = [("id", C.c_uint64), ("arrival_s", C.c_double), ("admit_s", C.c_double),
                ("done_s", C.c_double), ("iterations", C.c_uint32), ("batch", C.c_uint32),
                ("slot", C.c_uint32)]


class TuneResult(C.Structure):
    _fields_ = [("tile_items", C.c_uint32), ("batch_task_edges", C.c_uint32),
                ("bfs_push_threshold", C.c_double), ("pull_ms", C.c_double),
                ("batch_ms", C.c_double), ("bfs_ms", C.c_double),
                ("binned_src_bits", C.c_uint32), ("binned_dst_bits", C.c_uint32), ("binned_ms", C.c_double),
                ("batch_plan_packed", C.c_uint32), ("batch_plan_edges4", C.c_uint32),
                ("batch_plan_edges2", C.c_uint32), ("batch_plan_ms", C.c_double * 3)]


PE_TUNE_PULL, PE_TUNE_BATCH, PE_TUNE_BFS, PE_TUNE_BINNED = 1, 2, 4, 8

_vp = C.c_void_p
_u32, _u64, _i32, _dbl = C.c_uint32, C.c_uint64, C.c_int, C.c_double
_P = C.POINTER

# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    "pe_ctx_create": [_i32, _P(_vp)],
    "pe_ctx_destroy": [_vp],
    "pe_ctx_last_error": [_vp],
    "pe_ctx_sync": [_vp],
    "pe_ctx_mem_info": [_vp, _P(_u64), _P(_u64), _P(_u64)],
    "pe_ctx_timer_start": [_vp],
    "pe_ctx_timer_stop": [_vp, _P(_dbl)],
    "pe_ctx_set_kernel_timing": [_vp, _i32],
    "pe_ctx_kernel_time": [_vp, _i32, _P(_dbl), _P(_u64)],
    "pe_ctx_num_cus": [_vp, _P(_i32)],
    "pe_ctx_stream_rates": [_vp, _u64, _i32, _P(_dbl), _P(_dbl)],
    "pe_graph_create": [_vp, _vp, _vp, _vp, _u64, _u32, _u32, _P(_vp)],
    "pe_graph_create_csv": [_vp, C.c_char_p, _u64, _u32, _P(_vp)],
    "pe_graph_vertex_ids": [_vp, _vp],
    "pe_graph_create_reverse": [_vp, _u32, _P(_vp)],
    "pe_trustrank_select": [_vp, _vp, _vp, _vp, _vp, _vp],
    "pe_trustrank": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "pe_kcore": [_vp, C.c_int, _vp, _vp, _vp, _u32, _vp],
    "pe_connected_components": [_vp, _vp, _vp, _vp, _u32, _vp],
    "pe_nstep_degrees": [_vp, _vp, _u32, _u32, _i32, _vp, _vp, _P(_dbl)],
    "pe_hyperanf": [_vp, _vp, _vp, _u32, _vp, _vp, _vp, _vp],
    "pe_graph_create_rmat": [_vp, _u32, _u32, _u64, _dbl, _dbl, _dbl, _u32, _P(_vp)],
    "pe_rmat_edges": [_vp, _u32, _u64, _dbl, _dbl, _dbl, _u64, _u64, _vp, _vp, _vp],
    "pe_graph_destroy": [_vp],
    "pe_graph_get_info": [_vp, _P(GraphInfo)],
    "pe_graph_download_csc": [_vp, _vp, _vp, _vp],
    "pe_graph_download_csr": [_vp, _vp, _vp],
    "pe_graph_set_tile": [_vp, _u32],
    "pe_graph_set_pull_layout": [_vp, C.c_int, _vp],
    "pe_graph_permutation": [_vp, _vp],
    "pe_vec_create": [_vp, _i32, _u64, _P(_vp)],
    "pe_vec_destroy": [_vp],
    "pe_vec_fill": [_vp, _dbl],
    "pe_vec_copy": [_vp, _vp],
    "pe_vec_upload": [_vp, _vp, _u64],
    "pe_vec_download": [_vp, _vp, _u64],
    "pe_vec_download_f64": [_vp, _vp, _u64],
    "pe_vec_set": [_vp, _u64, _dbl],
    "pe_vec_sum": [_vp, _P(_dbl)],
    "pe_bitmap_create": [_vp, _u64, _P(_vp)],
    "pe_bitmap_destroy": [_vp],
    "pe_bitmap_clear": [_vp],
    "pe_bitmap_fill": [_vp],
    "pe_bitmap_set_bit": [_vp, _u64],
    "pe_bitmap_clr_bit": [_vp, _u64],
    "pe_bitmap_get_bit": [_vp, _u64, _P(_i32)],
    "pe_bitmap_count": [_vp, _P(_u64)],
    "pe_bitmap_msb": [_vp, _P(_u64)],
    "pe_bitmap_copy_from": [_vp, _vp],
    "pe_bitmap_upload": [_vp, _vp, _u64],
    "pe_bitmap_download": [_vp, _vp, _u64],
    "pe_bitmap_to_vector": [_vp, _vp, _u64, _P(_u64)],
    "pe_spread_message": [_vp, _i32, _vp, _vp, _vp, _P(_u64)],
    "pe_graph_out_degrees": [_vp, _vp],
    "pe_graph_in_degrees": [_vp, _vp],
    "pe_active_edges": [_vp, _vp, _P(_u64)],
    "pe_choose_mode": [_vp, _vp, _dbl, _P(_i32), _P(_u64)],
    "pe_pull_sum": [_vp, _vp, _vp, _i32],
    "pe_pull_sum_noalloc": [_vp, _vp, _vp, _i32],
    "pe_rank_apply": [_vp, _vp, _vp, _P(ApplyParams), _P(_dbl)],
    "pe_rank_iterate": [_vp, _vp, _vp, _i32, _P(ApplyParams), _P(_dbl)],
    "pe_pagerank": [_vp, _P(RankOpts), _vp, _vp, _P(RankStats)],
    "pe_ppr": [_vp, _vp, _u32, _P(RankOpts), _vp, _vp, _P(RankStats)],
    "pe_ppr_pushpull": [_vp, _u32, _P(RankOpts), _vp, _vp, _vp, _P(RankStats)],
    "pe_bfs": [_vp, _u32, _dbl, _vp, _vp, _vp, _vp, _u32, _P(BfsStats)],
    "pe_bfs_create": [_vp, _P(_vp)],
    "pe_bfs_destroy": [_vp],
    "pe_bfs_reset": [_vp, _u32, _vp],
    "pe_bfs_step": [_vp, _i32, _vp, _vp, _P(_u64), _P(_u64)],
    "pe_bfs_get_levels": [_vp, _vp],
    "pe_bfs_bind": [_vp, _vp],
    "pe_push_add": [_vp, _vp, _vp, _vp],
    "pe_rank_init": [_vp, _vp, _vp, _dbl, _i32],
    "pe_autotune": [_vp, _u32, _P(TuneResult)],
    "pe_autotune_save": [_vp, _P(TuneResult), C.c_char_p],
    "pe_autotune_load": [_vp, C.c_char_p, _P(TuneResult), _P(_i32)],
    "pe_ppr_batch_create": [_vp, _P(PprBatchOpts), _P(_vp)],
    "pe_ppr_batch_destroy": [_vp],
    "pe_ppr_batch_bytes": [_vp, _P(_u64)],
    "pe_ppr_batch_admit": [_vp, _u32, _vp, _u32],
    "pe_ppr_batch_release": [_vp, _u32],
    "pe_ppr_batch_step": [_vp, _u32],
    "pe_ppr_batch_poll": [_vp, _vp, _vp, _vp],
    "pe_ppr_batch_results": [_vp, _vp, _u32, _vp],
    "pe_ppr_batch_topk": [_vp, _vp, _u32, _u32, _vp, _vp],
    "pe_server_create": [_vp, _P(ServerOpts), _P(_vp)],
    "pe_server_destroy": [_vp],
    "pe_server_submit": [_vp, _u64, _vp, _u32, _dbl],
    "pe_server_pump": [_vp, _P(_u32)],
    "pe_server_take": [_vp, _u32, _vp, _vp, _vp, _P(_u32)],
    "pe_server_info": [_vp, _P(_u32), _P(_u32), _P(_u32), _P(_u64)],
    "pe_server_run_trace": [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _P(_dbl)],
}
_RESTYPE = {"pe_ctx_last_error": C.c_char_p}

_lib = None


def load():
    """Load libplato_hip.so; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m plato_amd.build`")
    L = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, C.c_int)
    _lib = L
    return L


def check(rc, ctx=None):
    if rc != PE_OK:
        msg = ""
        if ctx:
            raw = load().pe_ctx_last_error(ctx)
            msg = raw.decode() if raw else ""
        raise PlatoHipError(rc, msg)
