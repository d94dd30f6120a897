"""ctypes binding of libumaphip.so (the C ABI declared in include/umaphip.h).

The library is built in-tree by ``cuscan_amd/build.py`` (``__graft_entry__.build``).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libumaphip.so"

OK, INVALID_INPUT, INVALID_CONFIG, MEMORY, CAPACITY, NUMERICAL, BACKEND, OVERFLOW = range(8)


class UmapError(RuntimeError):
    """Raised for numerical, capacity, overflow and backend failures."""


class EpochArgs(C.Structure):
    _fields_ = [
        ("h", C.c_int64), ("n", C.c_int64), ("d", C.c_int64),
        ("head", C.c_void_p), ("head_s0", C.c_int64), ("head_s1", C.c_int64),
        ("fixed_tail", C.c_void_p), ("tail_s0", C.c_int64), ("tail_s1", C.c_int64),
        ("offsets", C.c_void_p), ("indices", C.c_void_p),
        ("source_order", C.c_void_p), ("negative_order", C.c_void_p),
        ("source_ranks", C.c_void_p),
        ("periods", C.c_void_p), ("next_positive", C.c_void_p),
        ("negative_periods", C.c_void_p), ("next_negative", C.c_void_p),
        ("updates", C.c_void_p), ("upd_s0", C.c_int64), ("upd_s1", C.c_int64),
        ("m", C.c_void_p), ("m_s0", C.c_int64), ("m_s1", C.c_int64),
        ("v", C.c_void_p), ("v_s0", C.c_int64), ("v_s1", C.c_int64),
        ("epoch", C.c_uint32),
        ("a", C.c_float), ("b", C.c_float), ("gamma", C.c_float), ("scale", C.c_float),
        ("negative_range", C.c_uint32),
        ("alpha", C.c_float), ("mu", C.c_float),
        ("correction1", C.c_float), ("correction2", C.c_float),
        ("momentum", C.c_int32), ("shared", C.c_int32), ("negatives", C.c_int32),
        ("exclude_graph_neighbors", C.c_int32),
        ("n_edges", C.c_int64),
    ]


class ClassicArgs(C.Structure):
    _fields_ = [
        ("h", C.c_int64), ("n", C.c_int64), ("d", C.c_int64), ("n_edges", C.c_int64),
        ("head", C.c_void_p), ("tail", C.c_void_p), ("tails", C.c_void_p), ("weights", C.c_void_p),
        ("epochs_per_sample", C.c_void_p), ("epoch_of_next_sample", C.c_void_p),
        ("epochs_per_negative_sample", C.c_void_p), ("epoch_of_next_negative_sample", C.c_void_p),
        ("rng_state", C.c_void_p), ("head_offsets", C.c_void_p), ("head_edges", C.c_void_p),
        ("mask", C.c_void_p), ("epoch", C.c_uint32),
        ("a", C.c_float), ("b", C.c_float), ("gamma", C.c_float), ("alpha", C.c_float),
        ("negative_sample_rate", C.c_float), ("move_other", C.c_int32), ("shared", C.c_int32),
    ]


class ServerParams(C.Structure):
    _fields_ = [
        ("n_neighbors", C.c_int32), ("epochs", C.c_int32),
        ("a", C.c_float), ("b", C.c_float), ("learning_rate", C.c_float),
        ("repulsion", C.c_float), ("local_connectivity", C.c_float),
        ("set_op_mix_ratio", C.c_float),
        ("negative_sample_rate", C.c_int64), ("negative_selection_range", C.c_int64),
        ("has_scale", C.c_int32), ("scale", C.c_float), ("adaptation_samples", C.c_int64),
        ("exclude_graph_neighbors", C.c_int32), ("transform_seed", C.c_uint64),
        ("max_batch_rows", C.c_int64), ("memory_bytes", C.c_int64),
    ]


class ServerCounters(C.Structure):
    _fields_ = [
        ("requests", C.c_int64), ("rows", C.c_int64), ("steps", C.c_int64),
        ("knn_batches", C.c_int64), ("mean_active", C.c_double), ("peak_active", C.c_int64),
        ("admission_waits", C.c_int64), ("memory_budget", C.c_int64),
        ("memory_in_use", C.c_int64), ("mean_queue_ms", C.c_double),
        ("mean_knn_ms", C.c_double), ("mean_graph_ms", C.c_double),
        ("mean_layout_ms", C.c_double),
    ]


class LayoutParams(C.Structure):
    _fields_ = [
        ("optimizer", C.c_int32), ("epochs", C.c_int32),
        ("a", C.c_float), ("b", C.c_float), ("repulsion", C.c_float),
        ("initial_alpha", C.c_float),
        ("negative_sample_rate", C.c_int64), ("negative_selection_range", C.c_int64),
        ("has_scale", C.c_int32), ("scale", C.c_float),
        ("adaptation_samples", C.c_int64),
        ("exclude_graph_neighbors", C.c_int32), ("good_initialization", C.c_int32),
        ("seed", C.c_uint64),
    ]


class LayoutStats(C.Structure):
    _fields_ = [
        ("epochs_run", C.c_int64), ("adaptations", C.c_int64), ("refreshes", C.c_int64),
        ("final_scale", C.c_float), ("device_ms", C.c_double),
        ("sweep_ms", C.c_double), ("sweep_launches", C.c_int64),
        ("refresh_ms", C.c_double), ("refresh_launches", C.c_int64),
        ("edge_balanced", C.c_int64), ("rank_ordered_epochs", C.c_int64),
        ("compact_records", C.c_int64), ("clock_schedule", C.c_int64),
    ]



class RecursiveParams(C.Structure):
    _fields_ = [
        ("coarsening_ratio", C.c_int64), ("base_init_threshold", C.c_int64),
        ("n_embedding_epochs", C.c_int64), ("hub_selection", C.c_int32),
        ("anchor_sample_size", C.c_int64),
    ]


class RecursiveLevel(C.Structure):
    _fields_ = [
        ("parts", C.c_int64), ("epochs", C.c_int64), ("a", C.c_float), ("b", C.c_float),
        ("good_initialization", C.c_int32), ("curve_a", C.c_double), ("curve_b", C.c_double),
    ]


class RecursiveStats(C.Structure):
    _fields_ = [
        ("depth", C.c_int32), ("rows", C.c_int64 * 32), ("edges", C.c_int64 * 32),
        ("sweeps", C.c_int32 * 32), ("host_ms", C.c_double), ("layout_ms", C.c_double),
        ("stage_ms", C.c_double * 8),
    ]

# Every exported symbol of include/umaphip.h, with its C signature.
_vp = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int32
SIGNATURES = {
    "umaphip_last_error": (C.c_char_p, []),
    "umaphip_version": (C.c_int32, []),
    "umaphip_device_count": (C.c_int32, []),
    "umaphip_modern_epoch": (C.c_int, [C.POINTER(EpochArgs), _vp]),
    "umaphip_debug_negatives": (C.c_int, [C.POINTER(EpochArgs), _i32, _vp, _vp, _vp, _vp]),
    "umaphip_project": (C.c_int, [_vp, _i64, _i64, _i64, _i64, _vp, _vp, _vp]),
    "umaphip_argsort_workspace_bytes": (C.c_int, [_i64, C.POINTER(C.c_size_t)]),
    "umaphip_stable_argsort_f32": (C.c_int, [_vp, _i64, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "umaphip_stable_argsort_f32_noalloc": (C.c_int, [_vp, _i64, _vp, _vp, _vp, _vp]),
    "umaphip_projection_ranks": (C.c_int, [_vp, _vp, _i64, _vp, _i64, _vp, _i32, _vp]),
    "umaphip_sample_resultants": (C.c_int, [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i64,
                                            C.c_uint32, C.c_float, C.c_float, C.c_float, _i32,
                                            _vp, _vp]),
    "umaphip_optimize_compact": (C.c_int, [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp,
                                           _vp, _vp, C.POINTER(LayoutParams), _i32,
                                           C.POINTER(LayoutStats)]),
    "umaphip_plan_create": (C.c_int, [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                                      _vp, C.POINTER(LayoutParams), _i32, C.POINTER(_vp)]),
    "umaphip_plan_run": (C.c_int, [_vp]),
    "umaphip_plan_download": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(LayoutStats)]),
    "umaphip_plan_stream": (_vp, [_vp]),
    "umaphip_plan_set_profiling": (C.c_int, [_vp, _i32]),
    "umaphip_plan_set_record_format": (C.c_int, [_vp, _i32]),
    "umaphip_plan_set_clock_schedule": (C.c_int, [_vp, _i32]),
    "umaphip_plan_set_trace": (C.c_int, [_vp, _i32, _i32]),
    "umaphip_plan_trace": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "umaphip_plan_destroy": (None, [_vp]),
    "umaphip_knn_index_create": (C.c_int, [_vp, _i32, _i64, _i64, _i64, C.POINTER(_vp)]),
    "umaphip_knn_query": (C.c_int, [_vp, _vp, _i64, _i32, _i32, _i64, _vp, _vp, _vp]),
    "umaphip_knn_index_destroy": (None, [_vp]),
    "umaphip_knn_candidates": (C.c_int, [_vp, _vp, _i64, _vp, _vp, _vp]),
    "umaphip_knn_set_rescan": (C.c_int, [_vp, _i32]),
    "umaphip_knn_certificate_stats": (C.c_int, [_vp, _vp, _vp, _vp]),
    "umaphip_fuzzy_graph": (C.c_int, [_vp, _vp, _i64, _i64, C.c_float, C.c_float, C.c_float, _i32,
                                      _i64, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(_i64)]),
    "umaphip_compact_graph": (C.c_int, [_i64, _i64, _i64, _vp, _vp, _vp, C.c_double, _i32,
                                        C.c_float, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                        C.POINTER(_i64)]),
    "umaphip_fit_curve": (C.c_int, [C.c_float, C.c_float, C.POINTER(C.c_float),
                                    C.POINTER(C.c_float)]),
    "umaphip_random_init": (C.c_int, [_i64, _i64, C.c_uint64, C.c_float, C.c_float, _i32, _vp]),
    "umaphip_transform_init": (C.c_int, [_vp, _vp, _i64, _i64, _vp, _i64, _i64, _vp]),
    "umaphip_fit_graph": (C.c_int, [_vp, _vp, _i64, _i64, C.c_float, C.c_float, C.c_float, _i32,
                                    C.c_float, C.c_double, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                    _vp]),
    "umaphip_transform_graph": (C.c_int, [_vp, _vp, _i64, _i64, _vp, _i64, _i64, C.c_float, _i32,
                                          C.c_float, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "umaphip_init_rescale": (C.c_int, [_vp, _i64, _i64, _vp, _vp, _vp, _i32, C.c_float,
                                       C.c_float]),
    "umaphip_recursive_init": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _i64, _i64, C.c_float,
                                         C.c_float, C.POINTER(RecursiveParams), C.c_uint64, _vp,
                                         C.POINTER(RecursiveStats)]),
    "umaphip_recursive_level_params": (C.c_int, [_i64, C.POINTER(RecursiveLevel), C.c_float,
                                                 C.c_float, C.POINTER(RecursiveParams),
                                                 C.POINTER(RecursiveLevel), C.POINTER(_i32)]),
    "umaphip_recursive_hubs": (C.c_int, [_i64, _vp, _vp, _vp, _i64, _i32, C.c_uint64, _vp]),
    "umaphip_recursive_propagate": (C.c_int, [_i64, _vp, _vp, _vp, _i32, _vp, C.POINTER(_i32)]),
    "umaphip_recursive_outliers": (C.c_int, [_i64, _vp, _vp, C.c_uint64, _vp]),
    "umaphip_recursive_remap": (C.c_int, [_vp, _i64, C.POINTER(_i64)]),
    "umaphip_recursive_aggregate": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "umaphip_recursive_transition": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "umaphip_recursive_anchor": (C.c_int, [_vp, _i64, _i64, _i64, _i64, C.c_uint64, _vp, _vp]),
    "umaphip_recursive_base_case": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _vp]),
    "umaphip_recursive_prolongate": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _vp]),
    "umaphip_recursive_procrustes": (C.c_int, [_vp, _i64, _vp, _vp, _i64]),
    "umaphip_classic_epoch": (C.c_int, [_vp, _vp]),
    "umaphip_plan_set_objective": (C.c_int, [_vp, _vp, _i32]),
    "umaphip_plan_objective": (C.c_int, [_vp, _vp, _i64, _vp]),
    "umaphip_cross_entropy_workspace_bytes": (C.c_int, [_i64, _vp]),
    "umaphip_cross_entropy": (C.c_int, [_vp, _i64, _vp, _i64, _i64, _vp, _vp, _vp, C.c_float,
                                        C.c_float, _vp, C.c_size_t, _vp, _vp]),
    "umaphip_server_create": (C.c_int, [_vp, _i64, _i64, _vp, _i64, _vp, _vp]),
    "umaphip_server_submit": (C.c_int, [_vp, _vp, _i64, _vp]),
    "umaphip_server_wait": (C.c_int, [_vp, _i64, C.c_double, _vp, _vp, _vp]),
    "umaphip_server_stats": (C.c_int, [_vp, _vp]),
    "umaphip_server_destroy": (None, [_vp]),
    "umaphip_batch_create": (C.c_int, [_i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64,
                                       _vp, _vp, _vp]),
    "umaphip_batch_tune": (C.c_int, [_vp, _i32, _vp, _vp]),
    "umaphip_batch_set_shape": (C.c_int, [_vp, _i32]),
    "umaphip_batch_run": (C.c_int, [_vp, _i32]),
    "umaphip_batch_set_persistent": (C.c_int, [_vp, _i32]),
    "umaphip_batch_set_clock_schedule": (C.c_int, [_vp, _i32]),
    "umaphip_batch_clock_schedule": (_i32, [_vp]),
    "umaphip_batch_download": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "umaphip_batch_destroy": (None, [_vp]),
    "umaphip_classic_optimize": (C.c_int, [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _i32,
                                           _i32, C.c_float, C.c_float, C.c_float, C.c_float, _i32,
                                           C.c_uint64, _i32, _vp, _vp, _vp]),
    "umaphip_orders_at_epoch": (C.c_int, [_i64, _i64, C.POINTER(LayoutParams), _i32, _i32, _vp,
                                          _vp]),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP extension first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status == OK:
        return
    msg = lib().umaphip_last_error().decode(errors="replace")
    if status in (INVALID_INPUT, INVALID_CONFIG):
        raise ValueError(msg)
    if status == MEMORY:
        raise MemoryError(msg)
    raise UmapError(msg)
