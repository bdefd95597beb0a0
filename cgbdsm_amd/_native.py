"""ctypes bindings for the two native libraries of the package.

libcgbdsm_host.so  host pipeline C ABI (include/cgbdsm_host.h)
libcgbdsm.so       MI355X Stage-2 engine C ABI (include/cgbdsm.h)

The structures mirror the C headers field by field.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
MAX_QV, MAX_QE = 12, 42
MAX_DIR = 2 * MAX_QE
MAX_BATCH_QUERIES = 256


class Plan(C.Structure):
    """cgbdsm_plan (include/cgbdsm.h)."""
    _fields_ = [
        ("qv", C.c_uint32), ("qe", C.c_uint32), ("ndir", C.c_uint32), ("ebits", C.c_uint32),
        ("eidx", C.c_uint8 * (MAX_QV * MAX_QV)),
        ("dir_to_edge", C.c_uint8 * MAX_DIR),
        ("qe_u", C.c_uint8 * MAX_QE), ("qe_v", C.c_uint8 * MAX_QE),
        ("go_vs", C.c_uint8 * MAX_QV), ("go_bni_offs", C.c_uint8 * (MAX_QV + 1)),
        ("go_bni", C.c_uint8 * MAX_DIR), ("go_cp", C.c_uint8 * MAX_QV), ("go_lcls", C.c_uint8 * MAX_QV),
        ("io_vs", (C.c_uint8 * MAX_QV) * MAX_QE),
        ("io_bni_offs", (C.c_uint8 * (MAX_QV + 1)) * MAX_QE),
        ("io_bni", (C.c_uint8 * MAX_DIR) * MAX_QE),
        ("rebuild_v", (C.c_uint8 * MAX_QV) * MAX_QE),
    ]


class Stats(C.Structure):
    """cgbdsm_stats (include/cgbdsm.h)."""
    _fields_ = [(n, C.c_uint64) for n in (
        "batches", "dv_count", "main_entries", "pool_used", "pool_cap", "last_ins", "last_del",
        "last_upd", "last_rows_touched", "last_idx_bytes", "last_frontier_rows",
        "last_frontier_depth", "last_matches", "compactions", "last_dfs_scanned", "last_dfs_probes",
        "last_dfs_donated", "gpma_rebuilds", "gpma_windows", "gpma_top_level", "last_dfs_steps",
        "last_dfs_step_lanes", "last_dfs_batched", "last_dfs_lds_cache")] + [
        (n, C.c_double) for n in ("ms_apply", "ms_update_index", "ms_masks", "ms_match", "ms_commit",
                                  "last_dfs_busy")]


class SplitStats(C.Structure):
    """cgbdsm_split_stats (include/cgbdsm.h)."""
    _fields_ = [("threads", C.c_int), ("ratio", C.c_double), ("batches", C.c_uint64),
                ("last_cpu_rows", C.c_uint64), ("last_gpu_rows", C.c_uint64), ("last_cpu_ms", C.c_double),
                ("last_gpu_ms", C.c_double), ("last_wait_ms", C.c_double), ("last_cpu_matches", C.c_uint64),
                ("mirror_entries", C.c_uint64),
                ("mirror_ms", C.c_double)]


class KTime(C.Structure):
    """cgbdsm_ktime (include/cgbdsm.h)."""
    _fields_ = [("ms", C.c_double), ("launches", C.c_uint64), ("bytes", C.c_uint64), ("items", C.c_uint64)]


KERNEL_CLASSES = ["sort", "merge_insert", "merge_remove", "update_index", "masks", "seed", "extend", "dfs",
                  "gpma_update", "gpma_extract"]


class SynthSpec(C.Structure):
    """cgh_synth_spec (include/cgbdsm_host.h)."""
    _fields_ = [
        ("scale", C.c_uint32), ("avg_degree", C.c_double),
        ("a", C.c_double), ("b", C.c_double), ("c", C.c_double),
        ("num_labels", C.c_uint32), ("zipf_s", C.c_double), ("initial_frac", C.c_double),
        ("stream_edges", C.c_uint64), ("seed", C.c_uint64),
        ("query_size", C.c_uint32), ("query_kind", C.c_uint32), ("query_seed", C.c_uint64),
    ]


class Batch(C.Structure):
    """cgh_batch (include/cgbdsm_host.h)."""
    _fields_ = [
        ("ins", C.POINTER(C.c_uint64)), ("n_ins", C.c_uint64),
        ("dele", C.POINTER(C.c_uint64)), ("n_del", C.c_uint64),
        ("upd", C.POINTER(C.c_uint64)), ("n_upd", C.c_uint64),
        ("edge_ok", C.POINTER(C.c_uint8)), ("dv", C.c_uint32), ("batch", C.c_uint32),
        ("pairs", C.c_uint64),
        ("ms_update_index", C.c_double), ("ms_construct_update", C.c_double),
        ("ms_prepare_all", C.c_double), ("ms_prepare_index", C.c_double),
    ]


def _lib_path(name):
    p = os.path.join(_HERE, name)
    if not os.path.exists(p):
        raise OSError(f"{name} is not built; run __graft_entry__.build() (expected at {p})")
    return p


_host = None
_dev = None


def host_lib():
    global _host
    if _host is None:
        L = C.CDLL(_lib_path("libcgbdsm_host.so"))
        P, U32, U64, I = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
        L.cgh_last_error.restype = C.c_char_p
        L.cgh_generate_files.argtypes = [C.POINTER(SynthSpec), C.c_char_p, C.c_char_p, C.c_char_p]
        L.cgh_session_open.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, U32, I, U32, C.POINTER(P)]
        L.cgh_session_open_synth.argtypes = [C.POINTER(SynthSpec), U32, I, U32, C.POINTER(P)]
        L.cgh_session_close.argtypes = [P]
        L.cgh_session_close.restype = None
        L.cgh_query_tables.argtypes = [P, C.POINTER(Plan)]
        L.cgh_initial_keys.argtypes = [P, C.POINTER(C.POINTER(U64)), C.POINTER(U64), C.POINTER(U32),
                                       C.POINTER(U32), C.POINTER(C.c_float)]
        L.cgh_make_plan.argtypes = [P, C.POINTER(U32), C.POINTER(C.c_float), C.POINTER(Plan)]
        L.cgh_num_batches.argtypes = [P]
        L.cgh_num_batches.restype = U32
        L.cgh_num_threads.argtypes = [P]
        L.cgh_num_threads.restype = U32
        L.cgh_next_batch.argtypes = [P, C.POINTER(Batch)]
        L.cgh_generate_binary.argtypes = [C.POINTER(SynthSpec), C.c_char_p, C.c_char_p]
        L.cgh_export_updates.argtypes = [P, C.c_void_p, U64, C.POINTER(U64)]
        L.cgh_next_batch_edges.argtypes = [P, U64, C.POINTER(Batch)]
        L.cgh_stream_left.argtypes = [P]
        L.cgh_stream_left.restype = U64
        L.cgh_old_ids.argtypes = [P, C.POINTER(C.POINTER(U32)), C.POINTER(U32)]
        L.cgh_export_graph.argtypes = [P, C.c_void_p, C.c_void_p, U64, C.c_void_p, C.POINTER(U32),
                                       C.POINTER(U64)]
        L.cgh_export_li.argtypes = [P, C.c_void_p, U32]
        L.cgh_export_cand.argtypes = [P, U32, U32, U32, C.c_void_p, U32, C.POINTER(C.c_int64)]
        L.cgh_query_info.argtypes = [P, C.c_void_p, C.c_void_p, C.POINTER(U32), C.POINTER(U32)]
        L.cgh_workload_open.argtypes = [C.POINTER(SynthSpec), C.POINTER(P)]
        L.cgh_workload_close.argtypes = [P]
        L.cgh_workload_close.restype = None
        L.cgh_workload_graph.argtypes = [P, C.POINTER(U32), C.POINTER(C.POINTER(C.c_int32)),
                                         C.POINTER(C.POINTER(U32)), C.POINTER(U64)]
        L.cgh_workload_stream.argtypes = [P, C.POINTER(C.POINTER(C.c_int32)), C.POINTER(U64)]
        L.cgh_workload_query.argtypes = [P, U32, U32, U64, C.c_void_p, C.c_void_p, C.POINTER(U32), C.POINTER(U32)]
        L.cgh_planner_open.argtypes = [C.c_void_p, U32, C.c_void_p, U32, U32, C.POINTER(P)]
        L.cgh_planner_close.argtypes = [P]
        L.cgh_planner_close.restype = None
        L.cgh_planner_tables.argtypes = [P, C.POINTER(Plan)]
        L.cgh_planner_make_plan.argtypes = [P, C.POINTER(U32), C.POINTER(C.c_float), C.POINTER(Plan)]
        _host = L
    return _host


def dev_lib():
    """The HIP engine; raises if it is not built (there is no CPU fallback)."""
    global _dev
    if _dev is None:
        # CGBDSM_DEV_LIB may name another build of the same library (tuning runs).
        L = C.CDLL(os.environ.get("CGBDSM_DEV_LIB") or _lib_path("libcgbdsm.so"))
        P, U32, U64, I = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
        L.cgbdsm_last_error.restype = C.c_char_p
        L.cgbdsm_version.restype = C.c_char_p
        L.cgbdsm_ctx_create.argtypes = [I, U64, C.POINTER(P)]
        L.cgbdsm_ctx_destroy.argtypes = [P]
        L.cgbdsm_ctx_memory.argtypes = [P, C.POINTER(U64), C.POINTER(U64)]
        L.cgbdsm_query_create.argtypes = [P, C.POINTER(Plan), C.POINTER(P)]
        L.cgbdsm_query_destroy.argtypes = [P]
        L.cgbdsm_query_set_plan.argtypes = [P, C.POINTER(Plan)]
        L.cgbdsm_index_init.argtypes = [P, C.c_void_p, U64, U32, C.c_void_p, C.c_void_p]
        L.cgbdsm_batch_stage.argtypes = [P, C.c_void_p, U64, C.c_void_p, U64, C.c_void_p, U64, C.c_void_p, U32]
        L.cgbdsm_transfer_update_all.argtypes = [P]
        L.cgbdsm_transfer_update_index.argtypes = [P, C.POINTER(I)]
        L.cgbdsm_build_local_index.argtypes = [P]
        L.cgbdsm_matching.argtypes = [P, C.POINTER(C.c_ulonglong), C.c_void_p]
        L.cgbdsm_matching_batch.argtypes = [C.POINTER(P), I, C.POINTER(C.c_ulonglong), C.POINTER(C.c_void_p)]
        L.cgbdsm_batch_update_global_index.argtypes = [P]
        L.cgbdsm_set_match_params.argtypes = [P, U64, U64, U32]
        L.cgbdsm_index_export.argtypes = [P, I, C.c_void_p, U64, C.POINTER(U64)]
        L.cgbdsm_index_export_static.argtypes = [P, I, C.c_void_p, U64, C.POINTER(U64)]
        L.cgbdsm_mask_export.argtypes = [P, U32, U32, C.c_void_p, U64, C.POINTER(I)]
        L.cgbdsm_match_tuples.argtypes = [P, C.c_void_p, U64, C.POINTER(U64), C.c_void_p]
        L.cgbdsm_stats_get.argtypes = [P, C.POINTER(Stats)]
        L.cgbdsm_run_batch.argtypes = [P, C.POINTER(C.c_ulonglong), C.c_void_p, C.POINTER(I)]
        L.cgbdsm_sync.argtypes = [P]
        L.cgbdsm_set_deadline_ms.argtypes = [P, C.c_double]
        L.cgbdsm_set_dfs_options.argtypes = [P, C.c_int, C.c_int]
        L.cgbdsm_set_dfs_tuning.argtypes = [P, U32, U32]
        L.cgbdsm_set_dfs_steal.argtypes = [P, U32, U32]
        L.cgbdsm_set_dfs_search.argtypes = [P, I]
        L.cgbdsm_build_frontier.argtypes = [P]
        L.cgbdsm_query_set_layout.argtypes = [P, C.c_int]
        L.cgbdsm_profile.argtypes = [P, I]
        L.cgbdsm_kernel_times.argtypes = [P, C.c_void_p, I]
        L.cgbdsm_batch_enqueue.argtypes = [P, C.c_void_p, U64, C.c_void_p, U64, C.c_void_p, U64, C.c_void_p, U32]
        L.cgbdsm_batch_stage_next.argtypes = [P, C.POINTER(I)]
        L.cgbdsm_graph_create.argtypes = [P, U32, C.c_void_p, C.c_void_p, U64, C.POINTER(P)]
        L.cgbdsm_graph_destroy.argtypes = [P]
        L.cgbdsm_graph_set_stream.argtypes = [P, C.c_void_p, U64]
        L.cgbdsm_graph_apply_range.argtypes = [P, U64, U64]
        L.cgbdsm_graph_info.argtypes = [P, C.POINTER(U64), C.POINTER(U64), C.POINTER(U64)]
        L.cgbdsm_calig_attach.argtypes = [P, P, C.c_void_p, C.c_void_p, C.c_void_p]
        L.cgbdsm_calig_update.argtypes = [P]
        L.cgbdsm_calig_match_deleted.argtypes = [P, C.POINTER(C.c_ulonglong), C.c_void_p]
        L.cgbdsm_calig_export_lit.argtypes = [P, C.c_void_p, U32]
        L.cgbdsm_calig_old_ids.argtypes = [P, C.c_void_p, U32, C.POINTER(U32)]
        L.cgbdsm_calig_stats.argtypes = [P, C.POINTER(C.c_double), C.c_void_p]
        L.cgbdsm_staged_keys.argtypes = [P, I, C.c_void_p, U64, C.POINTER(U64)]
        L.cgbdsm_set_cpu_split.argtypes = [P, I, C.c_double, C.c_double, U64]
        L.cgbdsm_cpu_split_stats.argtypes = [P, C.POINTER(SplitStats)]
        L.cgbdsm_cpu_mirror_export.argtypes = [P, I, C.c_void_p, U64, C.POINTER(U64)]
        _dev = L
    return _dev


DEV_SYMBOLS = [
    "cgbdsm_last_error", "cgbdsm_version", "cgbdsm_ctx_create", "cgbdsm_ctx_destroy", "cgbdsm_ctx_memory",
    "cgbdsm_query_create", "cgbdsm_query_destroy", "cgbdsm_query_set_plan", "cgbdsm_index_init",
    "cgbdsm_batch_stage", "cgbdsm_transfer_update_all", "cgbdsm_transfer_update_index",
    "cgbdsm_build_local_index", "cgbdsm_matching", "cgbdsm_matching_batch", "cgbdsm_batch_update_global_index",
    "cgbdsm_set_match_params", "cgbdsm_index_export", "cgbdsm_mask_export", "cgbdsm_match_tuples",
    "cgbdsm_stats_get", "cgbdsm_run_batch", "cgbdsm_sync", "cgbdsm_set_deadline_ms", "cgbdsm_set_dfs_options",
    "cgbdsm_profile", "cgbdsm_query_set_layout", "cgbdsm_set_dfs_tuning", "cgbdsm_build_frontier",
    "cgbdsm_kernel_times", "cgbdsm_batch_enqueue", "cgbdsm_batch_stage_next", "cgbdsm_index_export_static",
    "cgbdsm_graph_create", "cgbdsm_graph_destroy", "cgbdsm_graph_set_stream", "cgbdsm_graph_apply_range",
    "cgbdsm_graph_info", "cgbdsm_calig_attach", "cgbdsm_calig_update", "cgbdsm_calig_match_deleted",
    "cgbdsm_calig_export_lit",
    "cgbdsm_calig_old_ids", "cgbdsm_calig_stats", "cgbdsm_staged_keys", "cgbdsm_set_dfs_steal",
    "cgbdsm_set_dfs_search", "cgbdsm_set_cpu_split", "cgbdsm_cpu_split_stats", "cgbdsm_cpu_mirror_export",
]
