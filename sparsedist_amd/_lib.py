"""ctypes binding of libsd_hip.so (the C ABI declared in include/sd.h).

The library is built in-tree (``make -C sparsedist_amd``, or
``__graft_entry__.build()``).  There is no fallback: if the shared library is
missing or no GPU is visible, loading raises ``SparseDistError``.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsd_hip.so")

SD_OK, SD_EINVAL, SD_ENONFINITE, SD_ENOMEM, SD_ENOCAND, SD_ECAP, SD_ECANCELLED, \
    SD_EDEVICE, SD_EUNSUPPORTED, SD_EBUDGET = range(10)
SD_BINARY, SD_COSINE, SD_EUCLIDEAN, SD_MANHATTAN, SD_PEARSON, SD_COVARIANCE, SD_JS = range(7)
SD_PREC_AUTO, SD_PREC_EXACT, SD_PREC_FP32, SD_PREC_BF16 = range(4)

# every symbol include/sd.h declares
EXPORTS = (
    "sd_last_error", "sd_version", "sd_context_create", "sd_context_destroy", "sd_cancel",
    "sd_context_stats", "sd_context_set_timing", "sd_context_stream", "sd_matrix_upload_csc",
    "sd_matrix_free", "sd_matrix_drop_cache", "sd_knn", "sd_knn_cross", "sd_slab",
    "sd_topk_block", "sd_topk_rows_accum", "sd_snn_jaccard", "sd_snn", "sd_graph_info",
    "sd_graph_copy", "sd_graph_free", "sd_detect_cores", "sd_server_create", "sd_server_submit",
    "sd_server_wait", "sd_server_error", "sd_server_result", "sd_server_release",
    "sd_server_stats_get", "sd_server_destroy", "sd_autotune", "sd_server_resume",
    "sd_snn_jaccard_noalloc", "sd_context_set_buffer_budget", "sd_knn_kept",
    "sd_context_set_tuning", "sd_tune_samples",
)


class SparseDistError(RuntimeError):
    """An error reported by libsd_hip.so; ``code`` is the SD_* status."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class KnnParams(C.Structure):
    _fields_ = [("k", C.c_int32), ("method", C.c_int32), ("dist", C.c_int32),
                ("include_self", C.c_int32), ("drop_degenerate", C.c_int32),
                ("precision", C.c_int32), ("margin", C.c_int32), ("reserved", C.c_int32),
                ("block_size", C.c_int64)]


class Stats(C.Structure):
    _fields_ = [("slab_ms", C.c_double), ("slab_launches", C.c_int64),
                ("rerank_ms", C.c_double), ("rerank_launches", C.c_int64),
                ("snn_ms", C.c_double), ("snn_launches", C.c_int64),
                ("fallback_queries", C.c_int64), ("certified_queries", C.c_int64),
                ("escalated_queries", C.c_int64), ("overflow_queries", C.c_int64),
                ("filter_ms", C.c_double), ("filter_launches", C.c_int64),
                ("fused_calls", C.c_int64), ("exact_calls", C.c_int64),
                ("last_margin", C.c_int64), ("slab_products", C.c_int64),
                ("filter_pairs", C.c_int64), ("buffer_escalations", C.c_int64),
                ("escalation_list_queries", C.c_int64), ("query_blocks", C.c_int64),
                ("snn_device_inputs", C.c_int64), ("sample_stride", C.c_int64),
                ("sample_rank", C.c_int64)]


class TuneSample(C.Structure):
    _fields_ = [("method", C.c_int32), ("layout", C.c_int32), ("density_bucket", C.c_int32),
                ("row_bytes_log2", C.c_int32), ("list_len", C.c_int32), ("symmetric", C.c_int32),
                ("rule", C.c_int32), ("last_run", C.c_int32),
                ("stride", C.c_int32 * 4), ("rank", C.c_int32 * 4),
                ("ns_per_mpair", C.c_double * 4), ("runs", C.c_int64 * 4)]


class TuneInfo(C.Structure):
    _fields_ = [("bf16_ops", C.c_double), ("fp4_ops", C.c_double), ("merge_steps", C.c_double),
                ("bf16_fixed_ms", C.c_double), ("fp4_fixed_ms", C.c_double),
                ("exact_fixed_ms", C.c_double),
                ("calibrated", C.c_int32), ("reserved", C.c_int32),
                ("step_latency_us", C.c_double)]


class ServerParams(C.Structure):
    _fields_ = [("slots", C.c_int32), ("chunk_cols", C.c_int32), ("budget_bytes", C.c_int64),
                ("max_pending", C.c_int64), ("paused", C.c_int32), ("graphs", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("nrow", C.c_int64), ("ncol", C.c_int64), ("colptr", C.c_void_p),
                ("rowind", C.c_void_p), ("values", C.c_void_p), ("knn", KnnParams),
                ("snn", C.c_int32), ("borrow", C.c_int32), ("prune", C.c_double)]


class RequestInfo(C.Structure):
    _fields_ = [("done", C.c_int32), ("status", C.c_int32), ("k_eff", C.c_int32),
                ("warn", C.c_int32), ("queued_ms", C.c_double), ("run_ms", C.c_double),
                ("total_ms", C.c_double), ("finish_order", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("admitted", C.c_int64), ("completed", C.c_int64),
                ("refused", C.c_int64), ("pending", C.c_int64), ("active", C.c_int64),
                ("reserved_bytes", C.c_int64), ("query_cols", C.c_int64),
                ("packed_items", C.c_int64), ("waiting_callers", C.c_int64),
                ("graph_items", C.c_int64), ("graph_captures", C.c_int64),
                ("graph_fallbacks", C.c_int64), ("lane_admitted", C.c_int64)]


_lib = None
_lock = threading.Lock()


def load(path: str = LIB_PATH):
    """Load the shared library once; raise if it is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise SparseDistError(SD_EDEVICE, f"{path} is not built; run make -C sparsedist_amd")
        L = C.CDLL(path)
        vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        L.sd_last_error.restype = C.c_char_p
        L.sd_version.restype = C.c_char_p
        L.sd_context_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.sd_context_destroy.argtypes = [vp]
        L.sd_cancel.argtypes = [vp]
        L.sd_context_stats.argtypes = [vp, C.POINTER(Stats), C.c_int]
        L.sd_context_set_timing.argtypes = [vp, C.c_int]
        L.sd_context_set_buffer_budget.argtypes = [vp, C.c_int64]
        L.sd_knn_kept.argtypes = [vp, C.POINTER(vp), C.POINTER(i32), C.POINTER(i32),
                                  C.POINTER(i64)]
        L.sd_context_set_tuning.argtypes = [vp, i32]
        L.sd_tune_samples.argtypes = [C.POINTER(TuneSample), i32, C.POINTER(i32), i32]
        L.sd_context_stream.argtypes = [vp]
        L.sd_context_stream.restype = vp
        L.sd_matrix_upload_csc.argtypes = [vp, i64, i64, vp, vp, vp, C.POINTER(vp)]
        L.sd_matrix_free.argtypes = [vp]
        L.sd_matrix_drop_cache.argtypes = [vp]
        L.sd_knn.argtypes = [vp, vp, vp, C.POINTER(KnnParams), vp, vp, C.POINTER(i32),
                             C.POINTER(i32)]
        L.sd_knn_cross.argtypes = [vp, vp, vp, C.POINTER(KnnParams), vp, vp, C.POINTER(i32), vp,
                                   vp, C.POINTER(i32), C.POINTER(i32)]
        L.sd_slab.argtypes = [vp, i32, vp, vp, i32, vp]
        L.sd_topk_block.argtypes = [vp, vp, i64, i64, i32, i32, vp, vp, vp]
        L.sd_topk_rows_accum.argtypes = [vp, vp, i64, i64, i32, i32, vp, vp, i64, vp]
        L.sd_snn_jaccard.argtypes = [vp, vp, i32, i32, i32, vp]
        L.sd_snn_jaccard_noalloc.argtypes = [vp, vp, i32, i32, i32, vp]
        L.sd_snn.argtypes = [vp, vp, i32, i32, i32, dbl, i32, i32, i32, C.POINTER(vp)]
        L.sd_graph_info.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
        L.sd_graph_copy.argtypes = [vp, vp, vp, vp]
        L.sd_graph_free.argtypes = [vp]
        L.sd_detect_cores.argtypes = []
        L.sd_server_create.argtypes = [vp, vp, C.POINTER(ServerParams), C.POINTER(vp)]
        L.sd_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(i64)]
        L.sd_server_wait.argtypes = [vp, i64, dbl, C.POINTER(RequestInfo)]
        L.sd_server_error.argtypes = [vp, i64]
        L.sd_server_error.restype = C.c_char_p
        L.sd_server_result.argtypes = [vp, i64, vp, vp, C.POINTER(vp)]
        L.sd_server_release.argtypes = [vp, i64]
        L.sd_server_stats_get.argtypes = [vp, C.POINTER(ServerStats)]
        L.sd_server_destroy.argtypes = [vp]
        L.sd_server_resume.argtypes = [vp]
        L.sd_autotune.argtypes = [vp, i32, C.POINTER(TuneInfo)]
        _lib = L
        return L


def check(rc: int) -> None:
    if rc != SD_OK:
        raise SparseDistError(rc, load().sd_last_error().decode())
