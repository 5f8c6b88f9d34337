"""ctypes declarations of the C-ABI in include/gst.h (kept field-for-field in sync)."""
from __future__ import annotations

import ctypes as C
import os

GST_MAX_GROUPS = 12
GST_NOALLOC_SCRATCH_BYTES = 128 << 20
GST_NOALLOC_MAX_EDGES = 16384

GST_OK = 0
GST_INFEASIBLE = 1
GST_TOO_MANY_GROUPS = 2
GST_OOM = 3
GST_INVALID_ARG = 4
GST_HIP_ERROR = 5
GST_COST_OVERFLOW = 6
GST_NO_DEVICE = 7
GST_INTERNAL = 8

STATUS_NAMES = {
    0: "GST_OK", 1: "GST_INFEASIBLE", 2: "GST_TOO_MANY_GROUPS", 3: "GST_OOM", 4: "GST_INVALID_ARG",
    5: "GST_HIP_ERROR", 6: "GST_COST_OVERFLOW", 7: "GST_NO_DEVICE", 8: "GST_INTERNAL",
}

GST_W_INT32 = 0
GST_W_FLOAT32 = 1


class gst_csr(C.Structure):
    _fields_ = [("n", C.c_int32), ("m", C.c_int64), ("row_ptr", C.c_void_p), ("col", C.c_void_p),
                ("w", C.c_void_p), ("w_type", C.c_int32)]


class gst_groups(C.Structure):
    _fields_ = [("n_groups", C.c_int32), ("grp_ptr", C.c_void_p), ("grp_vtx", C.c_void_p)]


class gst_opts(C.Structure):
    _fields_ = [("ub", C.c_double), ("prune_lb", C.c_int32), ("want_tree", C.c_int32),
                ("rounds_per_sync", C.c_int32), ("chunk_arcs", C.c_int32), ("pull_permille", C.c_int32),
                ("grow_mode", C.c_int32), ("push_lanes", C.c_int32), ("half_bound", C.c_int32),
                ("merge_tile_kb", C.c_int32), ("auto_bound", C.c_int32), ("tour_bound", C.c_int32),
                ("reserved", C.c_int32 * 1)]


class gst_stats(C.Structure):
    _fields_ = [("grow_items", C.c_uint64), ("grow_arcs", C.c_uint64), ("grow_relax", C.c_uint64),
                ("grow_updates", C.c_uint64), ("grow_enqueued", C.c_uint64), ("merge_vertices", C.c_uint64),
                ("merge_pairs", C.c_uint64), ("grow_bytes", C.c_uint64), ("merge_bytes", C.c_uint64),
                ("grow_rounds", C.c_uint32), ("grow_launches", C.c_uint32), ("merge_launches", C.c_uint32),
                ("async_launches", C.c_uint32), ("t_total_ms", C.c_double), ("t_grow_ms", C.c_double),
                ("t_merge_ms", C.c_double), ("t_trace_ms", C.c_double), ("grow_new_states", C.c_uint64),
                ("grow_atomics", C.c_uint64), ("spider_cost", C.c_double), ("spider_tree_cost", C.c_double)]


class gst_result(C.Structure):
    _fields_ = [("cost", C.c_double), ("tree_cost", C.c_double), ("root", C.c_int32), ("n_edges", C.c_int32),
                ("eu", C.POINTER(C.c_int32)), ("ev", C.POINTER(C.c_int32)), ("ew", C.POINTER(C.c_double)),
                ("used_ub", C.c_uint8), ("fell_back", C.c_uint8), ("pad", C.c_uint8 * 2),
                ("n_groups", C.c_int32), ("stats", gst_stats)]


class gst_learned_stats(C.Structure):
    _fields_ = [("sub_vertices", C.c_uint32), ("sub_arcs", C.c_uint32), ("ub", C.c_double),
                ("t_score_select_ms", C.c_double), ("t_sub_solve_ms", C.c_double), ("t_full_solve_ms", C.c_double)]


class gst_batch_opts(C.Structure):
    _fields_ = [("slots", C.c_int32), ("rounds_per_step", C.c_int32), ("use_graph", C.c_int32),
                ("time_kernels", C.c_int32), ("max_steps", C.c_int32), ("reserved", C.c_int32 * 3)]


class gst_batch_stats(C.Structure):
    _fields_ = [("t_total_ms", C.c_double), ("queries_per_s", C.c_double), ("latency_p50_ms", C.c_double),
                ("latency_p99_ms", C.c_double), ("steps", C.c_uint64), ("grow_launches", C.c_uint64),
                ("grow_relax", C.c_uint64), ("grow_bytes", C.c_uint64), ("merge_bytes", C.c_uint64),
                ("queries", C.c_int64), ("slots", C.c_int32), ("rounds_per_step", C.c_int32),
                ("used_graph", C.c_int32), ("timed_kernels", C.c_int32), ("t_grow_ms", C.c_double),
                ("t_merge_ms", C.c_double), ("merge_launches", C.c_uint64), ("busy_slot_steps", C.c_uint64),
                ("reserved", C.c_int32 * 2)]


class gst_tune_report(C.Structure):
    _fields_ = [("chunk_arcs", C.c_int32), ("pull_permille", C.c_int32), ("push_lanes", C.c_int32),
                ("merge_tile_kb", C.c_int32), ("batch_rounds_per_step", C.c_int32), ("batch_push_lanes", C.c_int32),
                ("configs_tried", C.c_int32), ("pad", C.c_int32), ("default_ms", C.c_double), ("tuned_ms", C.c_double),
                ("batch_default_qps", C.c_double), ("batch_tuned_qps", C.c_double)]


GST_TUNE_BOUNDED = 1
GST_TUNE_BATCH = 2
GST_TUNE_EAGER = 4


class gst_host_csr(C.Structure):
    _fields_ = [("n", C.c_int32), ("m", C.c_int64), ("row_ptr", C.POINTER(C.c_int64)),
                ("col", C.POINTER(C.c_int32)), ("w", C.c_void_p), ("w_type", C.c_int32)]


# every function the header declares: name -> (restype, argtypes)
SIGNATURES = {
    "gst_ctx_create": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p)]),
    "gst_ctx_destroy": (None, [C.c_void_p]),
    "gst_last_error": (C.c_char_p, [C.c_void_p]),
    "gst_status_str": (C.c_char_p, [C.c_int]),
    "gst_graph_upload": (C.c_int, [C.c_void_p, C.POINTER(gst_csr), C.POINTER(C.c_void_p)]),
    "gst_graph_free": (None, [C.c_void_p]),
    "gst_graph_num_vertices": (C.c_int32, [C.c_void_p]),
    "gst_graph_num_arcs": (C.c_int64, [C.c_void_p]),
    "gst_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gst_groups), C.POINTER(gst_opts),
                            C.POINTER(gst_result)]),
    "gst_result_free": (None, [C.POINTER(gst_result)]),
    "gst_mst_postprocess": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "gst_solve_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gst_groups), C.POINTER(gst_opts), C.c_int32,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(gst_result)]),
    "gst_solve_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(gst_groups), C.POINTER(gst_opts),
                                  C.POINTER(gst_batch_opts), C.POINTER(gst_result), C.POINTER(C.c_int),
                                  C.POINTER(gst_batch_stats)]),
    "gst_validate_tree": (C.c_int, [C.POINTER(gst_csr), C.POINTER(gst_groups), C.c_int32, C.c_void_p,
                                    C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "gst_model_num_params": (C.c_int64, []),
    "gst_model_create": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "gst_model_free": (None, [C.c_void_p]),
    "gst_model_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gst_model_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gst_model_scores": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gst_groups), C.c_void_p]),
    "gst_model_loss_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gst_groups), C.c_void_p, C.c_float,
                                      C.POINTER(C.c_float), C.c_void_p]),
    "gst_model_adam_step": (C.c_int, [C.c_void_p, C.c_float]),
    "gst_solve_learned": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gst_groups), C.c_void_p, C.c_double,
                                    C.POINTER(gst_opts), C.POINTER(gst_result), C.POINTER(gst_learned_stats)]),
    "gst_solve_batch_learned": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(gst_groups), C.c_void_p,
                                          C.c_double, C.POINTER(gst_opts), C.POINTER(gst_batch_opts),
                                          C.POINTER(gst_result), C.POINTER(C.c_int), C.POINTER(gst_learned_stats),
                                          C.POINTER(gst_batch_stats)]),
    "gst_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(gst_groups), C.c_int32,
                               C.POINTER(gst_tune_report)]),
    "gst_host_csr_free": (None, [C.POINTER(gst_host_csr)]),
    "gst_gen_random_connected": (C.c_int, [C.c_int32, C.c_int64, C.c_int32, C.c_int32, C.c_uint64,
                                           C.POINTER(gst_host_csr)]),
    "gst_gen_barabasi_albert": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                          C.POINTER(gst_host_csr)]),
    "gst_gen_rmat": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32,
                               C.c_int32, C.c_double, C.c_double, C.c_uint64, C.POINTER(gst_host_csr)]),
    "gst_gen_groups": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.c_void_p, C.c_void_p]),
}

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgst.so")

_lib = None


def load() -> C.CDLL:
    """Load the in-tree libgst.so; fails loudly when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m gpu4gst_code_amd.build` "
                           "(or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
