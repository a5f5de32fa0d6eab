"""ctypes binding of the C ABI (include/sgm_c.h) exported by libsgm_hip.so.

The library is built in place by ``subgraphmatching_amd._build`` (or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there is no CPU
fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsgm_hip.so")

SGM_OK = 0
SGM_EINVAL = -1
SGM_EHIP = -2
SGM_ENOMEM = -3
SGM_EEMPTY = -4
SGM_EORDER = -5
SGM_ESTATE = -6
SGM_ESTOPPED = -7
SGM_EDEVICE = -8
SGM_ENODEV = -9

SGM_FILTER_NLF = 1
SGM_FILTER_GQL = 2
SGM_FILTER_CFL = 3
SGM_FILTER_CECI = 4
SGM_FILTER_DAF = 5

SGM_ENUM_CACHE_FILTER_FS = 0
SGM_ENUM_CACHE_FILTER = 1
SGM_ENUM_AUX_TABLES = 2
SGM_ENUM_AUX_TABLES_FS = 3
SGM_ENUM_DYN_DAF = 4
SGM_ENUM_DYN_VEQ = 5
SGM_ENUM_AUX_TABLES_GUPCD = 6
SGM_ENUM_AUX_TABLES_FS_GUPCD = 7

ERR_NAMES = {0: "SGM_OK", -1: "SGM_EINVAL", -2: "SGM_EHIP", -3: "SGM_ENOMEM", -4: "SGM_EEMPTY",
             -5: "SGM_EORDER", -6: "SGM_ESTATE", -7: "SGM_ESTOPPED", -8: "SGM_EDEVICE",
             -9: "SGM_ENODEV"}

# every symbol include/sgm_c.h declares (checked by tests/test_capi_symbols.py)
EXPORTS = [
    "sgm_version", "sgm_device_count", "sgm_ctx_create", "sgm_ctx_destroy", "sgm_ctx_last_error",
    "sgm_ctx_filter_ms",
    "sgm_stop", "sgm_graph_upload", "sgm_graph_destroy", "sgm_graph_max_degree", "sgm_query_create",
    "sgm_query_destroy", "sgm_filter", "sgm_filter_batch", "sgm_cand_counts", "sgm_cand_download",
    "sgm_cand_exist_download", "sgm_build_tables", "sgm_table_info", "sgm_table_download",
    "sgm_order_ri", "sgm_order_gql", "sgm_plan_export", "sgm_enumerate", "sgm_enumerate_batch",
    "sgm_batch_create", "sgm_batch_destroy", "sgm_batch_filter", "sgm_batch_order_gql", "sgm_batch_enumerate",
    "sgm_batch_enumerate_launch", "sgm_batch_enumerate_wait", "sgm_batch_last_ms", "sgm_batch_autotune",
    "sgm_costmodel_create", "sgm_costmodel_destroy", "sgm_costmodel_param_count", "sgm_costmodel_params",
    "sgm_costmodel_features", "sgm_costmodel_predict", "sgm_costmodel_grad", "sgm_costmodel_train",
    "sgm_costmodel_timing", "sgm_order_sample", "sgm_batch_rank_orders", "sgm_gql_work",
    "sgm_graph_set_labels", "sgm_query_set_labels", "sgm_plan_export_labeled", "sgm_batch_filter_labeled",
    "sgm_enumerate_noalloc", "sgm_order_qsi", "sgm_order_cfl", "sgm_order_ceci", "sgm_order_gup",
    "sgm_query_kcore", "sgm_query_bfs", "sgm_daf_start_vertex", "sgm_daf_weights", "sgm_set_daf_weights",
    "sgm_gup_guards",
]


SGM_ENUM_F_NO_BITMAPS = 1
SGM_ENUM_F_SERIAL_LAST = 2
SGM_ENUM_F_NO_SPLIT = 4
SGM_ENUM_F_SPLIT_ALL = 8


class EnumOpts(C.Structure):
    _fields_ = [("cap", C.c_uint64), ("method", C.c_int32), ("blocks", C.c_uint32),
                ("stack_words", C.c_uint32), ("flags", C.c_uint32), ("hub_min_degree", C.c_uint32),
                ("time_limit_ms", C.c_uint32), ("root_begin", C.c_uint32), ("root_end", C.c_uint32),
                ("split_chunks", C.c_uint32), ("reserved", C.c_uint32 * 1)]


class Stats(C.Structure):
    _fields_ = [("count", C.c_uint64), ("intersections", C.c_uint64), ("nodes", C.c_uint64 * 32),
                ("fs_prunes", C.c_uint64), ("elems_read", C.c_uint64), ("elems_out", C.c_uint64),
                ("elems_scanned", C.c_uint64), ("kernel_ms", C.c_double),
                ("capped", C.c_int32), ("stopped", C.c_int32), ("retries", C.c_uint32),
                ("timed_out", C.c_int32), ("busy_ns", C.c_uint64), ("probe_words", C.c_uint64)]


class CmSample(C.Structure):
    _fields_ = [("n", C.c_uint32), ("qoff", C.c_void_p), ("qadj", C.c_void_p), ("order", C.c_void_p),
                ("cand_count", C.c_void_p), ("cand_degsum", C.c_void_p)]


class SgmError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                          "(hipcc --offload-arch=gfx950)")
    L = C.CDLL(LIB_PATH)
    vp, u32p = C.c_void_p, C.POINTER(C.c_uint32)
    sig = {
        "sgm_version": (C.c_char_p, []),
        "sgm_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "sgm_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "sgm_ctx_destroy": (None, [vp]),
        "sgm_ctx_last_error": (C.c_char_p, [vp]),
        "sgm_ctx_filter_ms": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "sgm_stop": (C.c_int, [vp]),
        "sgm_graph_upload": (C.c_int, [vp, vp, vp, C.c_uint32, C.c_uint64, C.POINTER(vp)]),
        "sgm_graph_destroy": (None, [vp]),
        "sgm_graph_max_degree": (C.c_uint32, [vp]),
        "sgm_query_create": (C.c_int, [vp, vp, vp, C.c_uint32, C.POINTER(vp)]),
        "sgm_query_destroy": (None, [vp]),
        "sgm_filter": (C.c_int, [vp, C.c_int]),
        "sgm_filter_batch": (C.c_int, [vp, C.c_uint32, C.c_int, vp]),
        "sgm_cand_counts": (C.c_int, [vp, vp]),
        "sgm_cand_download": (C.c_int, [vp, C.c_uint32, vp, C.c_uint32]),
        "sgm_cand_exist_download": (C.c_int, [vp, vp, C.c_uint32]),
        "sgm_build_tables": (C.c_int, [vp]),
        "sgm_table_info": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64), u32p]),
        "sgm_table_download": (C.c_int, [vp, C.c_uint32, C.c_uint32, vp, vp, vp]),
        "sgm_order_ri": (C.c_int, [vp, vp, C.c_uint32, vp]),
        "sgm_order_gql": (C.c_int, [vp, vp]),
        "sgm_plan_export": (C.c_int, [vp, vp, C.c_uint32, vp, u32p, vp, vp, vp, vp, vp]),
        "sgm_plan_export_labeled": (C.c_int, [vp, vp, vp, C.c_uint32, vp, u32p, vp, vp, vp, vp, vp]),
        "sgm_graph_set_labels": (C.c_int, [vp, vp, C.c_uint32]),
        "sgm_query_set_labels": (C.c_int, [vp, vp, C.c_uint32]),
        "sgm_enumerate": (C.c_int, [vp, vp, C.POINTER(EnumOpts), C.POINTER(Stats)]),
        "sgm_enumerate_noalloc": (C.c_int, [vp, vp, C.POINTER(EnumOpts), C.POINTER(Stats)]),
        "sgm_enumerate_batch": (C.c_int, [vp, vp, C.c_uint32, C.POINTER(EnumOpts), vp]),
        "sgm_batch_create": (C.c_int, [vp, C.c_uint32, C.POINTER(vp)]),
        "sgm_batch_destroy": (None, [vp]),
        "sgm_batch_filter": (C.c_int, [vp, C.c_uint32, vp, vp, vp, C.c_int, vp, vp, vp]),
        "sgm_batch_filter_labeled": (C.c_int, [vp, C.c_uint32, vp, vp, vp, vp, C.c_int, vp, vp, vp]),
        "sgm_batch_order_gql": (C.c_int, [vp, C.c_uint32, vp]),
        "sgm_batch_enumerate": (C.c_int, [vp, vp, C.POINTER(EnumOpts), vp]),
        "sgm_batch_enumerate_launch": (C.c_int, [vp, vp, C.POINTER(EnumOpts)]),
        "sgm_batch_enumerate_wait": (C.c_int, [vp, vp]),
        "sgm_batch_last_ms": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "sgm_batch_autotune": (C.c_int, [vp, vp, C.POINTER(EnumOpts), C.POINTER(EnumOpts),
                                         C.POINTER(C.c_double)]),
        "sgm_costmodel_create": (C.c_int, [vp, C.c_uint64, C.POINTER(vp)]),
        "sgm_costmodel_destroy": (None, [vp]),
        "sgm_costmodel_param_count": (C.c_uint32, []),
        "sgm_costmodel_params": (C.c_int, [vp, vp, vp]),
        "sgm_costmodel_features": (C.c_int, [C.POINTER(CmSample), vp, vp, vp]),
        "sgm_costmodel_predict": (C.c_int, [vp, vp, C.c_uint32, vp]),
        "sgm_costmodel_grad": (C.c_int, [vp, vp, vp, C.c_uint32, vp, vp]),
        "sgm_costmodel_train": (C.c_int, [vp, vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64,
                                          vp]),
        "sgm_costmodel_timing": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                           C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "sgm_order_sample": (C.c_int, [vp, vp, C.c_uint32, C.c_uint64, vp]),
        "sgm_batch_rank_orders": (C.c_int, [vp, vp, C.c_uint32, C.c_uint64, C.c_float, vp, vp, vp, vp]),
        "sgm_gql_work": (C.c_int, [vp, vp]),
        "sgm_order_qsi": (C.c_int, [vp, vp]),
        "sgm_order_cfl": (C.c_int, [vp, vp]),
        "sgm_order_ceci": (C.c_int, [vp, vp]),
        "sgm_order_gup": (C.c_int, [vp, vp]),
        "sgm_query_kcore": (C.c_int, [vp, vp, C.c_uint32, vp]),
        "sgm_query_bfs": (C.c_int, [vp, vp, C.c_uint32, C.c_uint32, vp, vp, vp]),
        "sgm_daf_start_vertex": (C.c_int, [vp, vp]),
        "sgm_daf_weights": (C.c_int, [vp, vp, vp]),
        "sgm_set_daf_weights": (C.c_int, [vp, vp, vp]),
        "sgm_gup_guards": (C.c_int, [vp, vp, vp, vp, C.c_uint64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)
