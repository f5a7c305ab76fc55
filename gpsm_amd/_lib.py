"""ctypes bindings of libgpsm.so (the C ABI declared in include/gpsm.h and
include/gpsm_gen.h).

The library is built in-tree by ``make -C gpsm_amd``; importing this module
never compiles anything, and every entry point that needs the device fails
loudly when the library or the GPU is missing (there is no CPU fallback).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpsm.so")

MAX_QV = 64
FEAT = 32
HIDDEN = 128
N_PARAMS = HIDDEN * FEAT + HIDDEN + HIDDEN * HIDDEN + HIDDEN + HIDDEN + 1

ORDER_HEURISTIC, ORDER_LEARNED, ORDER_EXPLICIT = 0, 1, 2
Q_WALK_TREE, Q_WALK_INDUCED, Q_WALK_HALF, Q_CYCLE, Q_CLIQUE = 0, 1, 2, 3, 4

ERRORS = {
    0: "ok", -1: "invalid argument", -2: "out of memory", -3: "HIP runtime error", -4: "query too large",
    -5: "tuple capacity exceeded", -6: "no HIP device", -7: "query disconnected or without edges",
    -8: "malformed or unterminated input", -9: "join work budget exceeded",
    -10: "input larger than the fixed scratch of an allocation-free entry point",
}
EBUDGET = -9
ESCRATCH = -10
SCAN_NOALLOC_MAX = 32767 * 4096  # GPSM_SCAN_NOALLOC_MAX


class GpsmError(RuntimeError):
    def __init__(self, code, what):
        super().__init__("%s: %s (%d)" % (what, ERRORS.get(code, "error"), code))
        self.code = code


class CSR(C.Structure):
    _fields_ = [("n", C.c_uint32), ("adj", C.c_uint64), ("row_offset", C.POINTER(C.c_uint32)),
                ("col", C.POINTER(C.c_uint32)), ("labels", C.POINTER(C.c_uint32))]


class QSet(C.Structure):
    _fields_ = [("nq", C.c_uint32), ("v_off", C.POINTER(C.c_uint32)), ("labels", C.POINTER(C.c_uint32)),
                ("e_off", C.POINTER(C.c_uint32)), ("src", C.POINTER(C.c_uint32)), ("dst", C.POINTER(C.c_uint32))]


class QueryS(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint32), ("labels", C.c_void_p), ("src", C.c_void_p),
                ("dst", C.c_void_p)]


class Opts(C.Structure):
    _fields_ = [("order_mode", C.c_int32), ("explicit_order", C.c_void_p), ("want_tuples", C.c_uint32),
                ("tuple_cap", C.c_uint64), ("use_graph", C.c_uint32), ("block", C.c_uint32),
                ("join_block", C.c_uint32), ("join_cap", C.c_uint32), ("work_budget", C.c_uint64)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("q", C.c_uint32), ("count", C.c_uint64), ("hash", C.c_uint64),
                ("order", C.c_uint32 * MAX_QV), ("level_rows", C.c_uint64 * MAX_QV),
                ("cand", C.c_uint32 * MAX_QV), ("tuples", C.POINTER(C.c_uint32)), ("n_tuples", C.c_uint64),
                ("device_ms", C.c_double), ("join_work", C.c_uint64), ("level_work", C.c_uint64 * MAX_QV)]


class KStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint64), ("total_ms", C.c_double), ("bytes", C.c_double)]


_LIB = None

# name -> (restype, argtypes); every name here is exported by libgpsm.so
SIGNATURES = {
    "gpsm_init": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "gpsm_shutdown": (None, [C.c_void_p]),
    "gpsm_strerror": (C.c_char_p, [C.c_int]),
    "gpsm_device_name": (C.c_char_p, [C.c_void_p]),
    "gpsm_graph_upload": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.POINTER(C.c_void_p)]),
    "gpsm_graph_free": (C.c_int, [C.c_void_p]),
    "gpsm_graph_upload_edges": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.POINTER(C.c_void_p)]),
    "gpsm_graph_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]),
    "gpsm_graph_csr": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "gpsm_graph_device_bytes": (C.c_uint64, [C.c_void_p]),
    "gpsm_set_hub_policy": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64]),
    "gpsm_graph_hubs": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "gpsm_graph_set_directed": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gpsm_format_tuples": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p,
                                     C.c_uint64, C.POINTER(C.c_uint64)]),
    "gpsm_write_tuples": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p,
                                    C.POINTER(C.c_uint64)]),
    "gpsm_verify_directed": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.POINTER(C.c_uint64)]),
    "gpsm_match": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(QueryS), C.POINTER(Opts), C.POINTER(Result)]),
    "gpsm_match_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(QueryS), C.POINTER(Opts),
                                   C.POINTER(Result)]),
    "gpsm_result_free": (None, [C.POINTER(Result)]),
    "gpsm_filter": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(QueryS), C.c_void_p, C.c_void_p]),
    "gpsm_candidate_edges": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(QueryS), C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "gpsm_scan_u32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "gpsm_compact_bitmap": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.POINTER(C.c_uint64)]),
    "gpsm_mlp_create": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "gpsm_mlp_free": (None, [C.c_void_p]),
    "gpsm_mlp_num_params": (C.c_uint64, []),
    "gpsm_mlp_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gpsm_mlp_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gpsm_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "gpsm_mlp_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_float),
                                C.c_void_p]),
    "gpsm_mlp_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float,
                                      C.POINTER(C.c_float)]),
    "gpsm_set_planner": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gpsm_mlp_set_input_norm": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "gpsm_mlp_get_input_norm": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "gpsm_plan_features": (C.c_int, [C.c_void_p, C.POINTER(QueryS), C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p]),
    "gpsm_tuple_hash": (C.c_uint64, [C.c_void_p, C.c_uint32]),
    "gpsm_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "gpsm_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(QueryS), C.POINTER(Opts),
                                C.POINTER(C.c_double)]),
    "gpsm_tuning": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                              C.POINTER(C.c_uint32)]),
    "gpsm_set_tuning": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    "gpsm_profile_read": (C.c_int, [C.c_void_p, C.POINTER(KStat), C.c_uint32, C.POINTER(C.c_uint32)]),
    "gpsm_graph_cache_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint64)]),
    "gpsm_scan_u32_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    # generators and text I/O (include/gpsm_gen.h)
    "gpsm_gen_csr_from_edges": (C.c_int, [C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.POINTER(CSR)]),
    "gpsm_gen_er": (C.c_int, [C.c_uint32, C.c_uint64, C.c_uint64, C.POINTER(CSR)]),
    "gpsm_gen_chung_lu": (C.c_int, [C.c_uint32, C.c_uint64, C.c_double, C.c_uint64, C.POINTER(CSR)]),
    "gpsm_gen_rmat": (C.c_int, [C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                C.POINTER(CSR)]),
    "gpsm_gen_labels": (C.c_int, [C.POINTER(CSR), C.c_uint32, C.c_double, C.c_uint64]),
    "gpsm_gen_queries": (C.c_int, [C.POINTER(CSR), C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_uint64,
                                   C.c_uint64, C.POINTER(QSet)]),
    "gpsm_gen_free_csr": (None, [C.POINTER(CSR)]),
    "gpsm_gen_free_qset": (None, [C.POINTER(QSet)]),
    "gpsm_io_write_graph": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_uint32, C.c_void_p, C.c_uint64,
                                      C.c_void_p, C.c_void_p]),
    "gpsm_io_write_end": (C.c_int, [C.c_char_p]),
    "gpsm_io_count_graphs": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint64)]),
    "gpsm_io_write_csr": (C.c_int, [C.c_char_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gpsm_io_map_csr": (C.c_int, [C.c_char_p, C.POINTER(CSR)]),
    "gpsm_io_unmap_csr": (None, [C.POINTER(CSR)]),
}


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("gpsm_amd/libgpsm.so is not built: run `make -C gpsm_amd` "
                               "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc, what):
    if rc != 0:
        raise GpsmError(rc, what)
    return rc


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)
