"""ctypes binding of librayt_hip.so (the C ABI declared in include/rayt_hip.h).

Loading fails loudly when the library has not been built: there is no
fallback implementation of any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# RAYT_HIP_LIB selects another build of the same library (for comparing builds side by side).
LIB_PATH = os.environ.get("RAYT_HIP_LIB") or os.path.join(PKG_DIR, "librayt_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_DIR), "include", "rayt_hip.h")

RT_OK = 0
RT_ERR_RANGE = -6
RT_FLAG_SHADOWS = 1
RT_FLAG_COUNT = 2
RT_STATE_BYTES_PER_PIXEL = 48

P = C.c_void_p
I = C.c_int
U32 = C.c_uint32
U64 = C.c_uint64
F = C.c_float
D = C.c_double
FP = C.POINTER(C.c_float)
IP = C.POINTER(C.c_int)
U8P = C.POINTER(C.c_uint8)
U64P = C.POINTER(C.c_uint64)
PP = C.POINTER(C.c_void_p)


class TreeInfo(C.Structure):
    _fields_ = [("nodes_in_block", U32), ("blocks_count", U32), ("root_block_index", U32),
                ("channels_count", U32), ("bytes_in_block_content", U64), ("source_nodes", U64)]


class TerrainParams(C.Structure):
    _fields_ = [("depth", I), ("seed", U64), ("octaves", I), ("height_lo", F), ("height_hi", F),
                ("spheres", I), ("sphere_rmin", F), ("sphere_rmax", F), ("boxes", I),
                ("box_emin", F), ("box_emax", F), ("threads", I)]


# name -> (restype, argtypes)
SIGNATURES = {
    "rt_last_error": (C.c_char_p, []),
    "rt_version": (C.c_char_p, []),
    "rt_ctx_create": (I, [I, PP]),
    "rt_ctx_destroy": (I, [P]),
    "rt_ctx_finish": (I, [P]),
    "rt_ctx_compute_units": (I, [P, IP]),
    "rt_buf_create": (I, [P, U64, PP]),
    "rt_buf_destroy": (I, [P]),
    "rt_buf_size": (U64, [P]),
    "rt_buf_device_ptr": (P, [P]),
    "rt_buf_write": (I, [P, U64, U64, P, I, PP, I, PP]),
    "rt_buf_read": (I, [P, U64, U64, P, I, PP, I, PP]),
    "rt_buf_copy": (I, [P, U64, P, U64, U64, PP, I, PP]),
    "rt_buf_fill_u32": (I, [P, U64, U64, U32]),
    "rt_evt_wait": (I, [I, PP]),
    "rt_evt_retain": (I, [P]),
    "rt_evt_release": (I, [P]),
    "rt_evt_elapsed_ns": (I, [P, P, C.POINTER(D)]),
    "rt_evt_record": (I, [P, PP]),
    "rt_init_tracing_frame": (I, [P, P, I, I, FP, F]),
    "rt_raytracing_pass": (I, [P, P, P, I, I, P, P, I, P, P, I, I]),
    "rt_finish_tracing_frame": (I, [P, P, P, I, I, FP, P, P]),
    "rt_raytrace_depth": (I, [P, P, I, I, P, P, I, FP, F]),
    "rt_raytrace_restart": (I, [P, P, I, I, P, I, FP, F]),
    "rt_legacy_links_from_cache": (I, [P, P, P, P, U64]),
    "rt_render_plain": (I, [P, P, P, I, I, P, P, I, P, P, I, FP, F, FP, I, U64P]),
    "rt_plain_config": (I, [P, I, I, I]),
    "rt_node_to_block": (I, [P, I, I, P]),
    "rt_difference_flag": (I, [P, I, P, P]),
    "rt_compress": (I, [P, I, P, P, P]),
    "rt_sort_i32": (I, [P, P, I, P]),
    "rt_scan_i32": (I, [P, P, I, P]),
    "rt_sort_i32_noalloc": (I, [P, P, I]),
    "rt_merge_pass_i32": (I, [P, I, I, P, P]),
    "rt_scan_pass_i32": (I, [P, I, I, P, P]),
    "rt_tree_open": (I, [C.c_char_p, PP]),
    "rt_tree_build_sphere": (I, [F, F, F, F, I, I, PP]),
    "rt_tree_build_terrain": (I, [C.POINTER(TerrainParams), I, PP]),
    "rt_tree_build_terrain_file": (I, [C.POINTER(TerrainParams), I, C.c_char_p, PP]),
    "rt_tree_build_terrain_levelwise": (I, [C.POINTER(TerrainParams), I, PP]),
    "rt_tree_import_obj": (I, [C.c_char_p, I, I, I, PP]),
    "rt_tree_write": (I, [P, C.c_char_p]),
    "rt_tree_info_get": (I, [P, C.POINTER(TreeInfo)]),
    "rt_tree_channel": (I, [P, I, IP, C.c_char_p, I]),
    "rt_tree_load_block": (I, [P, U32, U8P, U8P, U64]),
    "rt_tree_destroy": (I, [P]),
    "rt_cache_create": (I, [P, P, I, PP]),
    "rt_cache_destroy": (I, [P]),
    "rt_cache_initial_fill": (I, [P]),
    "rt_cache_root_node_index": (I, [P]),
    "rt_cache_nodes_in_block": (I, [P]),
    "rt_cache_max_blocks": (I, [P]),
    "rt_cache_loaded_blocks": (I, [P]),
    "rt_cache_channel": (P, [P, I]),
    "rt_cache_channel_by_name": (P, [P, C.c_char_p]),
    "rt_cache_far_pointers": (P, [P]),
    "rt_cache_block_slot": (I, [P, I]),
    "rt_cache_slot_block": (I, [P, I]),
    "rt_cache_start_session": (I, [P]),
    "rt_cache_start_transaction": (I, [P]),
    "rt_cache_mark_block_used": (I, [P, I]),
    "rt_cache_mark_parent_used": (I, [P, I]),
    "rt_cache_request_block": (I, [P, I]),
    "rt_cache_transaction_filled": (I, [P]),
    "rt_cache_session_filled": (I, [P]),
    "rt_cache_end_transaction": (I, [P]),
    "rt_cache_end_session": (I, [P]),
    "rt_cache_lru_order": (I, [P, IP, I]),
    "rt_cache_counters": (I, [P, U64P]),
    "rt_feedback_create": (I, [P, I, PP]),
    "rt_feedback_destroy": (I, [P]),
    "rt_feedback_set_hits_to_extract": (I, [P, I]),
    "rt_feedback_set_seed": (I, [P, U64]),
    "rt_feedback_process": (I, [P, P, I, I]),
    "rt_feedback_list": (I, [P, I, IP, I]),
    "rt_feedback_stats": (I, [P, IP, IP]),
    "rt_tracer_create": (I, [P, I, I, PP]),
    "rt_tracer_destroy": (I, [P]),
    "rt_tracer_set_frame_time_limit": (I, [P, D]),
    "rt_tracer_set_lod_voxel_size": (I, [P, F]),
    "rt_tracer_set_max_passes": (I, [P, I]),
    "rt_tracer_set_flags": (I, [P, I]),
    "rt_tracer_set_seed": (I, [P, U64]),
    "rt_tracer_render_frame": (I, [P, FP]),
    "rt_tracer_out_image": (P, [P]),
    "rt_tracer_codes": (P, [P]),
    "rt_tracer_feedback": (P, [P]),
    "rt_tracer_frame_stats": (I, [P, U64P]),
    "rt_tracer_pass_faults": (I, [P, IP, I, IP, I]),
    "rt_tracer_timings": (I, [P, C.POINTER(D)]),
    "rt_tracer_set_pipelining": (I, [P, I]),
    "rt_tracer_trace_totals": (I, [P, C.POINTER(C.c_double), U64P]),
    "rt_camera_view_proj_inv": (I, [F, F, F, FP, F, F, F, FP]),
    "rt_batch_create": (I, [P, I, PP]),
    "rt_batch_destroy": (I, [P]),
    "rt_batch_config": (I, [P, I, I, I, I]),
    "rt_batch_render": (I, [P, C.c_void_p, I, P, P, U64P]),
    "rt_batch_last_ms": (I, [P, C.POINTER(D)]),
    "rt_batch_autotune": (I, [P, C.c_void_p, I, P, P, I, IP, C.POINTER(D)]),
    "rt_batch_add_scene": (I, [P, P]),
    "rt_server_create": (I, [P, C.c_void_p, PP]),
    "rt_server_destroy": (I, [P]),
    "rt_server_add_scene": (I, [P, P, I]),
    "rt_server_add_scene_pool": (I, [P, P, I, I]),
    "rt_server_submit": (I, [P, C.c_void_p, U64P]),
    "rt_server_submit_many": (I, [P, C.c_void_p, I, U64P]),
    "rt_server_poll": (I, [P, C.c_void_p, I, I]),
    "rt_server_drain": (I, [P, I]),
    "rt_server_release": (I, [P, U64]),
    "rt_server_stats_get": (I, [P, C.c_void_p]),
    "rt_server_reset_stats": (I, [P]),
}


class View(C.Structure):
    _fields_ = [("view_proj_inv", F * 16), ("width", I), ("height", I), ("lod_voxel_size", F), ("flags", I),
                ("scene", I), ("pixel_offset", U64)]


class ServerConfig(C.Structure):
    _fields_ = [("max_views_per_round", I), ("rounds_in_flight", I), ("framebuffer_bytes", U64),
                ("pool_bytes", U64), ("tile_shape", I), ("lds_levels", I), ("wgs_per_cu", I),
                ("max_block_loads", I)]


RT_OUT_IMAGE = 1
RT_OUT_CODES = 2
RT_DELIVER_COPY = 0
RT_DELIVER_PINNED = 1


class Request(C.Structure):
    _fields_ = [("view_proj_inv", F * 16), ("width", I), ("height", I), ("lod_voxel_size", F), ("flags", I),
                ("scene", I), ("outputs", I), ("delivery", I), ("pad", I), ("out_image", C.c_void_p),
                ("out_codes", C.c_void_p)]


class Completion(C.Structure):
    _fields_ = [("id", U64), ("status", I), ("round_views", I), ("submit_ms", D), ("admit_ms", D),
                ("done_ms", D), ("image", C.c_void_p), ("codes", C.c_void_p), ("fault_blocks", I), ("pad2", I)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", U64), ("completed", U64), ("failed", U64), ("rounds", U64), ("pixels", U64),
                ("primary_rays", U64), ("shadow_rays", U64), ("scene_loads", U64), ("scene_evictions", U64),
                ("max_rounds_in_flight", I), ("pad", I), ("p50_ms", D), ("p99_ms", D), ("mean_ms", D),
                ("max_ms", D), ("span_ms", D), ("mrays_per_s", D), ("mean_views_per_round", D),
                ("block_uploads", U64), ("load_transactions", U64)]

_lib = None


class RaytError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


def lib():
    """Loads librayt_hip.so once; raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() or python -m rayt_amd.build")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(fn_name: str, rc: int) -> int:
    if rc < 0:
        msg = lib().rt_last_error().decode(errors="replace")
        raise RaytError(fn_name, rc, msg)
    return rc


def call(name: str, *args) -> int:
    return check(name, getattr(lib(), name)(*args))


def declared_functions() -> list[str]:
    """Names of every function declared in include/rayt_hip.h."""
    import re
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(rt_[a-z0-9_]+)\s*\(", text)))
