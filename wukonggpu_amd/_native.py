"""ctypes bindings to the in-tree native libraries (built by wukonggpu_amd.build).

Loading fails loudly when a library is missing: there is no Python fallback
for any compute path.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")

# The small-query server is a persistent kernel on a stream of its own. HIP
# maps streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default), and
# work on a stream that shares the server's queue waits behind it. The
# library's streams (context, executors, served lanes, server) fit in 16:
# ask for that many unless the process set its own (the HIP runtime reads it
# when it starts).
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

# status codes (include/wukong_gpu.h)
WKG_OK, WKG_EINVAL, WKG_ENOMEM, WKG_EDEVICE, WKG_ERANGE, WKG_ESTATE = 0, -1, -2, -3, -4, -5
STEP_I2U, STEP_C2U, STEP_K2U, STEP_K2K, STEP_K2C = 0, 1, 2, 3, 4

u64, i64, u32, i32, dbl = C.c_uint64, C.c_int64, C.c_uint32, C.c_int32, C.c_double
P = C.POINTER
vp = C.c_void_p


class GenParams(C.Structure):
    _fields_ = [("seed", u64), ("num_entities", u64), ("num_preds", u32), ("num_types", u32),
                ("num_triples", u64), ("zipf_s", dbl), ("degree_cap", u64),
                ("type_sizes", P(u64)), ("num_hubs", u32), ("hub_degree", u64),
                ("num_linked", u32), ("link_frac", dbl), ("num_sibling", u32),
                ("sibling_src", P(u32)), ("sibling_dst", P(u32))]


class DatasetInfo(C.Structure):
    _fields_ = [("files", u64), ("triples", u64), ("index_ids", u64), ("max_index_id", u64),
                ("normal_ids", u64), ("max_normal_id", u64), ("has_str_index", i32), ("has_str_normal", i32)]


class BuildParams(C.Structure):
    _fields_ = [("main_load", dbl), ("num_threads", C.c_int)]


class PredMeta(C.Structure):
    _fields_ = [("partition_sz", u64), ("main_hdr_start", u64), ("main_hdr_end", u64),
                ("indrct_hdr_start", u64), ("indrct_hdr_end", u64), ("num_buckets", u64),
                ("edge_start", u64), ("edge_end", u64)]


class StepDesc(C.Structure):
    _fields_ = [("kind", i32), ("pid", i32), ("dir", i32), ("start_col", i32), ("end_col", i32),
                ("pad", i32), ("vid", i64), ("input", vp)]


class PlanStep(C.Structure):
    _fields_ = [("kind", i32), ("pid", i32), ("dir", i32), ("start_col", i32), ("end_col", i32),
                ("pad", i32), ("vid", i64)]


class BatchTuning(C.Structure):
    _fields_ = [("probe_wg_per_cu", i32), ("expand_wg_per_cu", i32), ("aux_wg_per_cu", i32),
                ("probe_tile_rows", i32), ("expand_tile_items", i32), ("schedule", i32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class BatchResult(C.Structure):
    _fields_ = [("offset", u64), ("rows", u64), ("cols", i32), ("status", i32), ("expanded", u64)]


class Result(C.Structure):
    _fields_ = [("table", P(i32)), ("rows", u64), ("cols", i32), ("status", i32),
                ("expanded", u64), ("latency_s", dbl)]


def _load(name):
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the native libraries first "
            "(python -m wukonggpu_amd.build or __graft_entry__.build())")
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


def _sig(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_gstore = None
_hip = None
_engine = None


def gstore_lib():
    global _gstore
    if _gstore is None:
        L = _load("libwkg_gstore.so")
        _sig(L, "wks_generate", C.c_int, P(GenParams), P(vp))
        _sig(L, "wks_triples_from_arrays", C.c_int, P(u32), P(u32), P(u32), u64, P(vp))
        _sig(L, "wks_triples_read_file", C.c_int, C.c_char_p, P(vp))
        _sig(L, "wks_triples_write_file", C.c_int, vp, C.c_char_p)
        _sig(L, "wks_dataset_read", C.c_int, C.c_char_p, P(vp), P(DatasetInfo))
        _sig(L, "wks_dataset_write", C.c_int, vp, C.c_char_p, C.c_int)
        _sig(L, "wks_triples_count", u64, vp)
        for n in ("s", "p", "o"):
            _sig(L, f"wks_triples_{n}", P(u32), vp)
        _sig(L, "wks_triples_free", None, vp)
        _sig(L, "wks_build", C.c_int, vp, P(BuildParams), P(vp))
        _sig(L, "wks_gstore_free", None, vp)
        for n in ("num_slots", "num_buckets", "num_buckets_ext", "num_keys", "num_edges"):
            _sig(L, f"wks_{n}", u64, vp)
        _sig(L, "wks_slots", P(u64), vp)
        _sig(L, "wks_edges", P(u32), vp)
        _sig(L, "wks_num_metas", C.c_int, vp)
        _sig(L, "wks_metas", P(PredMeta), vp)
        _sig(L, "wks_get_edges", C.c_int, vp, u64, u32, C.c_int, P(P(u32)), P(u64))
        _sig(L, "wks_gstore_save", C.c_int, vp, C.c_char_p)
        _sig(L, "wks_gstore_load", C.c_int, C.c_char_p, P(vp))
        _sig(L, "wks_last_error", C.c_char_p)
        _gstore = L
    return _gstore


def hip_lib():
    global _hip
    if _hip is None:
        L = _load("libwkg_hip.so")
        _sig(L, "wkg_init", C.c_int, C.c_int, u64, P(vp))
        _sig(L, "wkg_destroy", C.c_int, vp)
        _sig(L, "wkg_last_error", C.c_char_p, vp)
        _sig(L, "wkg_load_graph", C.c_int, vp, P(u64), u64, P(u32), u64, P(PredMeta), C.c_int, u64, u32)
        _sig(L, "wkg_ensure_predicate", C.c_int, vp, C.c_int)
        _sig(L, "wkg_predicate_resident", C.c_int, vp, C.c_int)
        _sig(L, "wkg_prefetch_predicate", C.c_int, vp, C.c_int, C.c_int)
        _sig(L, "wkg_preload_predicates", C.c_int, vp, P(i32), C.c_int)
        _sig(L, "wkg_graph_prefetches", C.c_int, vp, P(u64))
        _sig(L, "wkg_graph_stats", C.c_int, vp, P(u64), P(u64), P(u64))
        _sig(L, "wkg_table_from_host", C.c_int, vp, P(i32), u64, C.c_int, vp, P(vp))
        _sig(L, "wkg_table_to_host", C.c_int, vp, vp, P(i32), vp)
        _sig(L, "wkg_table_shape", C.c_int, vp, P(u64), P(C.c_int))
        _sig(L, "wkg_table_device_ptr", vp, vp)
        _sig(L, "wkg_table_wrap_device", C.c_int, vp, vp, u64, C.c_int, P(vp))
        _sig(L, "wkg_table_free", C.c_int, vp, vp)
        _sig(L, "wkg_table_slice", C.c_int, vp, vp, u64, u64, P(vp))
        _sig(L, "wkg_table_split", C.c_int, vp, vp, C.c_int, C.c_int, vp, P(vp), P(u64), P(u64))
        _sig(L, "wkg_table_concat", C.c_int, vp, C.c_int, P(vp), vp, P(vp))
        _sig(L, "wkg_set_small_steps", C.c_int, vp, C.c_int)
        _sig(L, "wkg_thread_launch_rows", C.c_int, vp, u64)
        _sig(L, "wkg_set_table_limit", C.c_int, vp, u64)
        _sig(L, "wkg_set_probe_tile_rows", C.c_int, vp, C.c_int)
        _sig(L, "wkg_set_lineage", C.c_int, vp, C.c_int)
        _sig(L, "wkg_set_group_filter_rows", C.c_int, vp, C.c_uint64)
        _sig(L, "wkg_build_graph", C.c_int, vp, P(C.c_uint32), P(C.c_uint32), P(C.c_uint32), u64, dbl, P(dbl))
        _sig(L, "wkg_graph_shape", C.c_int, vp, P(u64), P(u64), P(C.c_int))
        _sig(L, "wkg_graph_download", C.c_int, vp, vp, vp, vp)
        _sig(L, "wkg_pool_stats", C.c_int, vp, P(u64), P(u64), P(u64), P(u64))
        _sig(L, "wkg_step_i2u", C.c_int, vp, C.c_int, C.c_int, vp, P(vp), P(u64))
        _sig(L, "wkg_step_c2u", C.c_int, vp, i64, C.c_int, C.c_int, vp, P(vp), P(u64))
        _sig(L, "wkg_step_k2u", C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, vp, P(vp), P(u64))
        _sig(L, "wkg_step_k2u_noalloc", C.c_int, vp, vp, u64, C.c_int, C.c_int, C.c_int, C.c_int, vp, u64, vp,
             P(u64), P(u64))
        _sig(L, "wkg_step_k2k", C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, P(vp), P(u64))
        _sig(L, "wkg_step_k2c", C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, i64, vp, P(vp), P(u64))
        _sig(L, "wkg_query_small", C.c_int, vp, P(PlanStep), C.c_int, vp, P(i32), u64, P(u64), P(C.c_int), P(u64))
        _sig(L, "wkg_query_small_submit", C.c_int, vp, P(PlanStep), C.c_int, P(u64))
        _sig(L, "wkg_step_k2u_filter", C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
             i64, vp, P(vp), P(C.c_int), P(u64), P(u64))
        _sig(L, "wkg_query_small_collect", C.c_int, vp, u64, C.c_int, P(i32), u64, P(u64), P(C.c_int), P(u64))
        _sig(L, "wkg_set_small_query_server", C.c_int, vp, C.c_int)
        _sig(L, "wkg_set_fanout_blocks", C.c_int, vp, C.c_int, u64)
        _sig(L, "wkg_thread_context_begin", C.c_int, vp)
        _sig(L, "wkg_thread_context_end", C.c_int, vp)
        _sig(L, "wkg_set_small_query_server_times", C.c_int, vp, C.c_double, C.c_double)
        _sig(L, "wkg_step_batch", C.c_int, vp, C.c_int, P(StepDesc), vp, P(vp), P(u64))
        _sig(L, "wkg_profile_enable", C.c_int, vp, C.c_int)
        _sig(L, "wkg_profile_read", C.c_int, vp, C.c_char_p, P(u64), P(dbl), P(dbl))
        _sig(L, "wkg_bandwidth_probe", C.c_int, vp, u64, C.c_int, P(dbl))
        _sig(L, "wkg_random_access_probe", C.c_int, vp, u64, C.c_int, P(dbl))
        _sig(L, "wkg_batch_create", C.c_int, vp, C.c_int, C.c_int, u64, vp, P(vp))
        _sig(L, "wkg_batch_destroy", C.c_int, vp)
        _sig(L, "wkg_batch_set_tuning", C.c_int, vp, P(BatchTuning))
        _sig(L, "wkg_batch_get_tuning", C.c_int, vp, P(BatchTuning))
        _sig(L, "wkg_batch_launch", C.c_int, vp, C.c_int, P(PlanStep), P(i32))
        _sig(L, "wkg_batch_busy", C.c_int, vp)
        _sig(L, "wkg_batch_collect", C.c_int, vp, P(BatchResult), P(P(i32)), P(u64))
        _sig(L, "wkg_batch_run", C.c_int, vp, C.c_int, P(PlanStep), P(i32), P(BatchResult), P(P(i32)), P(u64))
        _sig(L, "wkg_batch_stats", C.c_int, vp, P(u64), P(u64))
        _sig(L, "wkg_batch_reserve_results", C.c_int, vp, u64)
        _sig(L, "wkg_batch_autotune", C.c_int, vp, C.c_int, P(PlanStep), P(i32), C.c_int, P(BatchTuning), P(dbl))
        _hip = L
    return _hip


def engine_lib():
    global _engine
    if _engine is None:
        hip_lib()
        L = _load("libwkg_engine.so")
        _sig(L, "wke_create", C.c_int, vp, P(vp))
        _sig(L, "wke_destroy", C.c_int, vp)
        _sig(L, "wke_last_error", C.c_char_p, vp)
        _sig(L, "wke_free", None, vp)
        _sig(L, "wke_run_query", C.c_int, vp, P(i64), C.c_int, P(Result))
        _sig(L, "wke_run_plain", C.c_int, vp, P(i64), P(i64), C.c_int, P(Result))
        _sig(L, "wke_run_batched", C.c_int, vp, P(i64), P(i64), C.c_int, P(Result))
        _sig(L, "wke_run_rounds", C.c_int, vp, P(i64), P(i64), C.c_int, P(Result))
        _sig(L, "wke_run_batched_packed", C.c_int, vp, P(i64), P(i64), C.c_int, P(Result), P(u64),
             P(P(i32)), P(u64))
        _sig(L, "wke_configure", C.c_int, vp, C.c_int, C.c_int, u64)
        _sig(L, "wke_stats", C.c_int, vp, P(u64))
        _sig(L, "wke_set_small_queries", C.c_int, vp, C.c_int)
        _sig(L, "wke_autotune", C.c_int, vp, P(i64), P(i64), C.c_int, C.c_int, C.c_int, P(BatchTuning), P(dbl))
        _sig(L, "wke_set_tuning", C.c_int, vp, P(BatchTuning), C.c_int)
        _sig(L, "wke_set_heavy_mode", C.c_int, vp, C.c_int)
        _sig(L, "wke_serve_stats", C.c_int, vp, P(u64), P(u64), P(u64), P(u64))
        _sig(L, "wke_set_serve_small", C.c_int, vp, C.c_int)
        _sig(L, "wke_set_fused_filters", C.c_int, vp, C.c_int)
        _sig(L, "wke_set_pipeline", C.c_int, vp, C.c_int)
        _sig(L, "wke_set_heavy_workers", C.c_int, vp, C.c_int)
        _sig(L, "wke_serve_small_stats", C.c_int, vp, P(u64), P(u64))
        _sig(L, "wke_serve", C.c_int, vp, P(i64), P(i64), C.c_int, P(dbl), P(C.c_uint8), u64, C.c_int,
             P(Result), P(dbl))
        _sig(L, "wke_serve_closed", C.c_int, vp, P(i64), P(i64), C.c_int, P(C.c_uint8), C.c_int, C.c_int, dbl,
             dbl, u64, C.c_int, P(C.c_uint8), P(Result), P(dbl), P(dbl))
        _engine = L
    return _engine
