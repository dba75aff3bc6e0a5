"""ctypes binding of libgazelle_amd.so (the C ABI declared in include/gazelle_amd.h).

The library must have been built (``python -m gazelle_plugin_amd.build`` or
``__graft_entry__.build()``); there is no CPU fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgazelle_amd.so")

GZ_OK = 0
STATUS_NAMES = {1: "OutOfMemory", 2: "KeyError", 3: "TypeError", 4: "Invalid", 5: "IOError",
                6: "CapacityError", 7: "IndexError", 9: "UnknownError", 10: "NotImplemented",
                100: "HipError"}


class GazelleError(RuntimeError):
    """A non-OK status from the C ABI; ``code`` is the arrow::StatusCode value."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code
        self.msg = msg


class OutOfMemory(GazelleError):
    pass


class Invalid(GazelleError):
    pass


class NotImplementedStatus(GazelleError):
    pass


_CLASSES = {1: OutOfMemory, 4: Invalid, 10: NotImplementedStatus}


class GzBatch(ctypes.Structure):
    _fields_ = [("num_rows", ctypes.c_int64), ("buf_addrs", ctypes.POINTER(ctypes.c_int64)),
                ("buf_sizes", ctypes.POINTER(ctypes.c_int64)), ("num_bufs", ctypes.c_int32)]


class GzSplitOptions(ctypes.Structure):
    _fields_ = [("buffer_size", ctypes.c_int32), ("prefer_spill", ctypes.c_int32),
                ("offheap_per_task", ctypes.c_int64), ("tile_rows", ctypes.c_int32),
                ("hash_seed", ctypes.c_int32), ("append_row_ids", ctypes.c_int32),
                ("batch_compress_threshold", ctypes.c_int32)]


class GzSplitMetrics(ctypes.Structure):
    _fields_ = [("total_compute_pid_time", ctypes.c_int64), ("total_write_time", ctypes.c_int64),
                ("total_spill_time", ctypes.c_int64), ("total_compress_time", ctypes.c_int64),
                ("total_bytes_written", ctypes.c_int64), ("total_bytes_spilled", ctypes.c_int64)]


class GzAggAction(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("input0", ctypes.c_int32), ("input1", ctypes.c_int32),
                ("nan_check", ctypes.c_int32)]


class GzAggOptions(ctypes.Structure):
    _fields_ = [("placement", ctypes.c_int32), ("tile_rows", ctypes.c_int32),
                ("load_factor", ctypes.c_double), ("expected_groups", ctypes.c_int64)]


class GzAggPlan(ctypes.Structure):
    _fields_ = [("num_cols", ctypes.c_int32), ("col_types", ctypes.POINTER(ctypes.c_int32)),
                ("num_actions", ctypes.c_int32), ("actions", ctypes.POINTER(GzAggAction))]


class GzSplitPlan(ctypes.Structure):
    _fields_ = [("num_cols", ctypes.c_int32), ("col_types", ctypes.POINTER(ctypes.c_int32)),
                ("num_partitions", ctypes.c_int32), ("key_cols", ctypes.POINTER(ctypes.c_int32)),
                ("num_keys", ctypes.c_int32)]


class GzProbePlan(ctypes.Structure):
    _fields_ = [("num_probe_cols", ctypes.c_int32), ("probe_types", ctypes.POINTER(ctypes.c_int32)),
                ("probe_key_col", ctypes.c_int32), ("join_type", ctypes.c_int32),
                ("num_result_cols", ctypes.c_int32), ("result_sides", ctypes.POINTER(ctypes.c_int32)),
                ("result_indices", ctypes.POINTER(ctypes.c_int32))]


class GzFragment(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("relation", ctypes.c_int32), ("num_batches", ctypes.c_int32),
                ("batches", ctypes.POINTER(GzBatch)), ("arrival_s", ctypes.c_double)]


class GzFragmentStats(ctypes.Structure):
    _fields_ = [("arrival_s", ctypes.c_double), ("admitted_s", ctypes.c_double), ("done_s", ctypes.c_double),
                ("rows", ctypes.c_int64), ("out_rows", ctypes.c_int64), ("status", ctypes.c_int32),
                ("slot", ctypes.c_int32)]


_lib = None

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_PI32 = ctypes.POINTER(ctypes.c_int32)
_PI64 = ctypes.POINTER(ctypes.c_int64)
_PP = ctypes.POINTER(ctypes.c_void_p)

_SIGS = {
    "gz_last_error": (ctypes.c_char_p, []),
    "gz_context_create": (ctypes.c_int, [ctypes.c_int, _I64, _PP]),
    "gz_context_destroy": (ctypes.c_int, [_P]),
    "gz_context_set_stream": (ctypes.c_int, [_P, _P]),
    "gz_context_stream": (_P, [_P]),
    "gz_context_synchronize": (ctypes.c_int, [_P]),
    "gz_context_bytes_in_use": (_I64, [_P]),
    "gz_context_peak_bytes": (_I64, [_P]),
    "gz_context_set_profiling": (ctypes.c_int, [_P, ctypes.c_int]),
    "gz_context_kernel_time": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), _PI64]),
    "gz_context_reset_kernel_times": (ctypes.c_int, [_P]),
    "gz_split_options_defaults": (None, [ctypes.POINTER(GzSplitOptions)]),
    "gz_split_make": (ctypes.c_int, [_P, ctypes.c_char_p, _I32, _PI32, _I32, _PI32, _I32,
                                     ctypes.POINTER(GzSplitOptions), _PP]),
    "gz_split_batch": (ctypes.c_int, [_P, _I64, _PI64, _PI64, _I32]),
    "gz_split_batches": (ctypes.c_int, [_P, _I32, ctypes.POINTER(GzBatch)]),
    "gz_split_stop": (ctypes.c_int, [_P]),
    "gz_split_partition_rows": (ctypes.c_int, [_P, _PI64]),
    "gz_split_num_partitions": (ctypes.c_int, [_P]),
    "gz_split_num_output_columns": (ctypes.c_int, [_P]),
    "gz_split_num_segments": (ctypes.c_int, [_P]),
    "gz_split_segment_rows": (ctypes.c_int, [_P, _I32, _PI64]),
    "gz_split_segment_column": (ctypes.c_int, [_P, _I32, _I32, _PP, _PP, _PP, _PP, _PI64, _PI64]),
    "gz_split_segment_batch_counts": (ctypes.c_int, [_P, _I32, _PI32, _PI32]),
    "gz_split_compute_pid_time": (_I64, [_P]),
    "gz_split_raw_partition_lengths": (ctypes.c_int, [_P, _PI64]),
    "gz_split_batch_rows": (ctypes.c_int, [_P, _PI64]),
    "gz_split_get_metrics": (ctypes.c_int, [_P, ctypes.POINTER(GzSplitMetrics)]),
    "gz_split_release": (ctypes.c_int, [_P]),
    "gz_split_ipc_prepare": (ctypes.c_int, [_P, ctypes.c_void_p, _I64, _I32, _PI64, _PI64]),
    "gz_split_ipc_write": (ctypes.c_int, [_P, ctypes.c_void_p, _I64]),
    "gz_ipc_record_batch_message": (ctypes.c_int, [_I64, _I32, _PI64, _I32, _PI64, _I64, ctypes.c_void_p, _I64,
                                                   _PI64]),
    "gz_split_set_compress_type": (ctypes.c_int, [_P, _I32]),
    "gz_ipc_read_batches": (ctypes.c_int, [ctypes.c_void_p, _I64, _I64, _I64, _I32, _PI32, _I32, _PI64, _PI64,
                                           _PI64, _PI32, _PI64]),
    "gz_ipc_decompress": (ctypes.c_int, [_P, ctypes.c_void_p, _I64, _I64, _I32, _PI32, ctypes.c_void_p, _I64]),
    "gz_agg_result_bitmap": (ctypes.c_int, [_P, _I32, ctypes.POINTER(ctypes.c_void_p)]),
    "gz_agg_set_result_bitmap": (ctypes.c_int, [_P, _I32, _I32, ctypes.POINTER(ctypes.c_void_p)]),
    "gz_probe_result_bitmap": (ctypes.c_int, [_P, _I32, ctypes.POINTER(ctypes.c_void_p)]),
    "gz_validity_bitmap": (ctypes.c_int, [_P, ctypes.c_void_p, _I64, ctypes.c_void_p]),
    "gz_ipc_decompress_noalloc": (ctypes.c_int, [_P, ctypes.c_void_p, _I64, _I64, _I32, _PI32, ctypes.c_void_p, _I64]),
}

# symbols added by later operators are bound when present
_OPTIONAL_SIGS = {
    "gz_agg_options_defaults": (None, [ctypes.POINTER(GzAggOptions)]),
    "gz_agg_make": (ctypes.c_int, [_P, _I32, _PI32, _I32, ctypes.POINTER(GzAggAction),
                                   ctypes.POINTER(GzAggOptions), _PP]),
    "gz_agg_process": (ctypes.c_int, [_P, _I64, _PI64, _PI64, _I32]),
    "gz_agg_process_batches": (ctypes.c_int, [_P, _I32, ctypes.POINTER(GzBatch)]),
    "gz_join_agg_make": (ctypes.c_int, [_P, _P, _I32, _PI32, _I32, _I32, _PI32, _PI32, _I32,
                                        ctypes.POINTER(GzAggAction), ctypes.POINTER(GzAggOptions), _PP]),
    "gz_join_agg_process_batches": (ctypes.c_int, [_P, _I32, ctypes.POINTER(GzBatch)]),
    "gz_agg_num_groups": (ctypes.c_int, [_P, _PI64]),
    "gz_agg_num_result_columns": (ctypes.c_int, [_P]),
    "gz_agg_result_type": (ctypes.c_int, [_P, _I32]),
    "gz_agg_finish": (ctypes.c_int, [_P]),
    "gz_agg_result_column": (ctypes.c_int, [_P, _I32, _PP, _PP]),
    "gz_agg_result_binary": (ctypes.c_int, [_P, _I32, _PP, _PP, _PP]),
    "gz_agg_release": (ctypes.c_int, [_P]),
    "gz_agg_set_make": (ctypes.c_int, [_P, _I32, _PI32, _I32, ctypes.POINTER(GzAggAction),
                                       ctypes.POINTER(GzAggOptions), _I32, _PP]),
    "gz_agg_set_process": (ctypes.c_int, [_P, _I32, ctypes.POINTER(GzBatch), _PI32]),
    "gz_agg_set_num_groups": (ctypes.c_int, [_P, _I32, _PI64]),
    "gz_agg_set_result_column": (ctypes.c_int, [_P, _I32, _I32, _PP, _PP]),
    "gz_agg_set_placement": (ctypes.c_int, [_P, _I32, _PI32]),
    "gz_agg_set_prepare": (ctypes.c_int, [_P, _I32, ctypes.POINTER(GzBatch), _PI32]),
    "gz_agg_set_run": (ctypes.c_int, [_P]),
    "gz_agg_set_capture": (ctypes.c_int, [_P]),
    "gz_agg_set_launch": (ctypes.c_int, [_P]),
    "gz_agg_set_autotune": (ctypes.c_int, [_P, _I32, ctypes.POINTER(GzBatch), _PI32, _PI32]),
    "gz_server_create": (ctypes.c_int, [_I32, _I64, _I32, ctypes.POINTER(GzAggPlan), ctypes.POINTER(GzSplitPlan),
                                        ctypes.POINTER(GzProbePlan), ctypes.POINTER(GzAggPlan), _PP]),
    "gz_server_add_relation": (ctypes.c_int, [_P, _P, _PI32]),
    "gz_server_run": (ctypes.c_int, [_P, _I32, ctypes.POINTER(GzFragment), _I32, ctypes.POINTER(GzFragmentStats)]),
    "gz_server_fragment_agg": (ctypes.c_int, [_P, _I32, _PP]),
    "gz_server_fragment_split": (ctypes.c_int, [_P, _I32, _PP]),
    "gz_server_release_results": (ctypes.c_int, [_P]),
    "gz_server_destroy": (ctypes.c_int, [_P]),
    "gz_relation_make": (ctypes.c_int, [_P, _I32, _PI32, _I32, _I32, _PP]),
    "gz_relation_make_keys": (ctypes.c_int, [_P, _I32, _PI32, _I32, _PI32, _I32, _PP]),
    "gz_relation_normalize_keys": (ctypes.c_int, [_P, _I32, _PI32]),
    "gz_relation_append": (ctypes.c_int, [_P, _I64, _PI64, _PI64, _I32]),
    "gz_relation_append_batches": (ctypes.c_int, [_P, _I32, ctypes.POINTER(GzBatch)]),
    "gz_relation_finish": (ctypes.c_int, [_P]),
    "gz_relation_num_keys": (ctypes.c_int, [_P, _PI64]),
    "gz_relation_has_null_key": (ctypes.c_int, [_P, _PI32]),
    "gz_relation_release": (ctypes.c_int, [_P]),
    "gz_probe_make": (ctypes.c_int, [_P, _I32, _PI32, _I32, _I32, _I32, _I32, _PI32, _PI32, _PP]),
    "gz_probe_make_keys": (ctypes.c_int, [_P, _I32, _PI32, _I32, _PI32, _I32, _I32, _I32, _PI32, _PI32, _PP]),
    "gz_probe_process": (ctypes.c_int, [_P, _I64, _PI64, _PI64, _I32, _PI64]),
    "gz_probe_process_batches": (ctypes.c_int, [_P, _I32, ctypes.POINTER(GzBatch), _PI64, _PI64]),
    "gz_probe_result_column": (ctypes.c_int, [_P, _I32, _PP, _PP, _PP, _PP]),
    "gz_probe_set_context": (ctypes.c_int, [_P, _P]),
    "gz_probe_release": (ctypes.c_int, [_P]),
}


def _bind_process_hip_runtime():
    """torch ships its own libamdhip64.so (SONAME libamdhip64.so.7).  Importing
    torch first makes the dynamic loader bind this library's libamdhip64.so.7
    dependency to that already-mapped runtime, so one HIP runtime serves the
    process; loading ours first would map a second runtime that torch cannot
    share the device with."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Loads the HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        _bind_process_hip_runtime()
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        for name, (res, args) in _OPTIONAL_SIGS.items():
            f = getattr(L, name, None)
            if f is not None:
                f.restype = res
                f.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    if rc != GZ_OK:
        msg = lib().gz_last_error()
        msg = msg.decode() if msg else ""
        raise _CLASSES.get(rc, GazelleError)(rc, msg)


def i32_array(vals):
    vals = list(vals)
    return (ctypes.c_int32 * max(1, len(vals)))(*vals)


def i64_array(vals):
    vals = list(vals)
    return (ctypes.c_int64 * max(1, len(vals)))(*vals)
