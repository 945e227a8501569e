"""ctypes binding of libhv.so (the C ABI declared in include/hv/hv.h).

The library is loaded from this package directory only. There is no fallback:
if libhv.so is missing or was built without a symbol, loading fails loudly.
"""
from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libhv.so"

u32 = C.c_uint32
i32 = C.c_int32
u64 = C.c_uint64
f32 = C.c_float
vp = C.c_void_p
P = C.POINTER

# ---------------------------------------------------------------------------
# ctypes mirrors of the public structs (layout checked against numpy dtypes in
# helio_amd.layouts and against the C static_asserts).
# ---------------------------------------------------------------------------


class CullConfig(C.Structure):
    _fields_ = [
        ("lod_count", u32),
        ("lod_distances", f32 * 7),
        ("flags", u32),
        ("screen_width", u32),
        ("screen_height", u32),
    ]


class CullOutputs(C.Structure):
    _fields_ = [
        ("ids", vp),
        ("args", vp),
        ("stats", vp),
        ("bucket_count", u32),
        ("id_capacity", u32),
    ]


class ViewRequest(C.Structure):
    _fields_ = [
        ("scene", vp),
        ("camera", vp),
        ("first", u32),
        ("count", u32),
        ("hiz", vp),
    ]


class VgFrameData(C.Structure):
    _fields_ = [
        ("meshlets", vp),
        ("objects", vp),
        ("instances", vp),
        ("work_items", vp),
        ("meshlet_count", u32),
        ("object_count", u32),
        ("work_item_count", u32),
        ("max_draw_count", u32),
        ("buffer_version", u64),
        ("instance_version", u64),
        ("instance_dirty_start", u32),
        ("instance_dirty_count", u32),
    ]


class VgOutputs(C.Structure):
    _fields_ = [("indirect", vp), ("metadata", vp), ("counters", vp), ("capacity", u32)]


class DrawBuffers(C.Structure):
    _fields_ = [
        ("instances", vp),
        ("aabbs", vp),
        ("draw_calls", vp),
        ("indirect", vp),
        ("stats", vp),
        ("draw_count", u32),
        ("instance_capacity", u32),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("cull", CullConfig),
        ("max_batch_views", u32),
        ("max_inflight", u32),
        ("max_batch_candidates", u64),
        ("device_budget_bytes", u64),
        ("max_buckets", u32),
        ("use_graph", u32),
        ("copy_ids", u32),
        ("max_vg_draws", u32),
        ("grow_pools", u32),
        ("device_headroom_bytes", u64),
    ]


class VgRequest(C.Structure):
    _fields_ = [
        ("scene", vp),
        ("width", u32), ("height", u32),
        ("lod_error_threshold_px", f32),
        ("flags", u32),
        ("hiz", vp),
        ("capacity", u32),
        ("indirect_out", vp), ("metadata_out", vp), ("counters_out", vp),
    ]


class ServerRequest(C.Structure):
    _fields_ = [
        ("view", ViewRequest),
        ("ids_out", vp),
        ("args_out", vp),
        ("stats_out", vp),
        ("vg", VgRequest),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", u64), ("completed", u64), ("rejected", u64), ("batches", u64),
        ("latency_p50_us", C.c_double), ("latency_p99_us", C.c_double), ("latency_max_us", C.c_double),
        ("latency_mean_us", C.c_double), ("mean_batch_views", C.c_double),
        ("device_bytes_peak", u64), ("device_bytes_slots", u64), ("device_bytes_free", u64), ("pool_grows", u64),
        ("memory_waits", u64),
    ]


class ShadowMatrixBuffers(C.Structure):
    _fields_ = [
        ("lights", vp), ("shadow_matrices", vp), ("camera", vp), ("shadow_dirty", vp), ("shadow_hashes", vp),
        ("light_count", u32), ("matrix_capacity", u32), ("slot_capacity", u32),
    ]


class ShadowCullBuffers(C.Structure):
    _fields_ = [
        ("shadow_matrices", vp), ("instances", vp), ("src_indirect", vp), ("face_dirty", vp),
        ("draw_count", u32), ("face_count", u32), ("instance_capacity", u32),
    ]


class RebuildInputs(C.Structure):
    _fields_ = [
        ("instances", vp), ("aabbs", vp), ("draws", vp), ("objects", vp), ("materials", vp),
        ("object_count", u32), ("material_count", u32), ("group_hidden", C.c_uint64),
    ]


class RebuildOutputs(C.Structure):
    _fields_ = [
        ("instances", vp), ("aabbs", vp), ("draw_calls", vp), ("indirect", vp), ("visibility", vp),
        ("gpu_slots", vp), ("ranges", vp), ("shadow_static", vp), ("shadow_movable", vp), ("counts", vp),
        ("scene", vp),
    ]


class TuneResult(C.Structure):
    _fields_ = [("wg", u32), ("ipt", u32), ("classify_ms", f32), ("total_ms", f32)]


# name -> (restype, argtypes)
_SIGS: dict[str, tuple] = {
    "hv_last_error": (C.c_char_p, []),
    "hv_ctx_create": (i32, [i32, P(vp)]),
    "hv_ctx_destroy": (None, [vp]),
    "hv_ctx_stream": (i32, [vp, P(vp)]),
    "hv_ctx_synchronize": (i32, [vp]),
    "hv_ctx_device_info": (i32, [vp, P(u32), P(u64)]),
    "hv_device_alloc": (i32, [vp, u64, P(vp)]),
    "hv_device_free": (i32, [vp, vp]),
    "hv_memcpy_h2d": (i32, [vp, vp, vp, u64, vp]),
    "hv_memcpy_d2d": (i32, [vp, vp, vp, u64, vp]),
    "hv_memcpy_d2h": (i32, [vp, vp, vp, u64, vp]),
    "hv_memset_d32": (i32, [vp, vp, u32, u64, vp]),
    "hv_timer_record": (i32, [vp, u32, vp]),
    "hv_timer_elapsed_ms": (i32, [vp, u32, u32, P(f32)]),
    "hv_extract_frustum_planes": (i32, [vp, vp]),
    "hv_mip_levels": (u32, [u32, u32]),
    "hv_instance_cull_data_from_instances": (i32, [vp, u32, vp]),
    "hv_scene_create": (i32, [vp, u32, P(vp)]),
    "hv_scene_destroy": (None, [vp]),
    "hv_scene_upload": (i32, [vp, vp, vp, u32, u32, vp]),
    "hv_scene_upload_soa": (i32, [vp, vp, vp, vp, u32, u32, vp]),
    "hv_scene_set_mesh_lods": (i32, [vp, vp, u32, u32]),
    "hv_scene_upload_device": (i32, [vp, vp, u32, u32, vp]),
    "hv_stream_create": (i32, [vp, P(vp)]),
    "hv_stream_destroy": (i32, [vp, vp]),
    "hv_stream_synchronize": (i32, [vp]),
    "hv_scene_swap_remove": (i32, [vp, u32, vp]),
    "hv_scene_truncate": (i32, [vp, u32]),
    "hv_scene_read_soa": (i32, [vp, vp, vp, u32, u32]),
    "hv_scene_info": (i32, [vp, P(u32), P(u32), P(u32), P(u64)]),
    "hv_hiz_create": (i32, [vp, u32, u32, u32, P(vp)]),
    "hv_hiz_destroy": (None, [vp]),
    "hv_hiz_bind_depth": (i32, [vp, vp]),
    "hv_hiz_upload_depth": (i32, [vp, vp, vp]),
    "hv_hiz_prepare": (i32, [vp, u64]),
    "hv_hiz_execute": (i32, [vp, vp]),
    "hv_hiz_levels": (i32, [vp, P(u32)]),
    "hv_hiz_level_size": (i32, [vp, u32, P(u32), P(u32)]),
    "hv_hiz_read_level": (i32, [vp, u32, vp, u64]),
    "hv_inst_cull_create": (i32, [vp, P(CullConfig), P(vp)]),
    "hv_inst_cull_destroy": (None, [vp]),
    "hv_inst_cull_prepare": (i32, [vp, vp, vp, u32, u32, vp]),
    "hv_inst_cull_execute": (i32, [vp, vp]),
    "hv_inst_cull_outputs": (i32, [vp, P(CullOutputs)]),
    "hv_inst_cull_read": (i32, [vp, vp, u32, P(u32), vp, u32, vp]),
    "hv_inst_cull_execute_timed": (i32, [vp, vp, P(f32)]),
    "hv_inst_cull_set_shape": (i32, [vp, u32, u32]),
    "hv_inst_cull_noalloc": (i32, [vp, P(CullConfig), vp, vp, u32, u32, vp, vp, vp, vp, vp]),
    "hv_batch_create": (i32, [vp, P(CullConfig), u32, P(vp)]),
    "hv_batch_destroy": (None, [vp]),
    "hv_batch_set_shape": (i32, [vp, u32, u32]),
    "hv_batch_prepare": (i32, [vp, P(ViewRequest), u32]),
    "hv_batch_execute": (i32, [vp, vp, u32]),
    "hv_batch_execute_timed": (i32, [vp, vp, P(f32)]),
    "hv_batch_outputs": (i32, [vp, u32, P(CullOutputs)]),
    "hv_batch_read": (i32, [vp, u32, vp, u32, P(u32), vp, u32, vp]),
    "hv_batch_wait": (i32, [vp]),
    "hv_batch_device_bytes": (i32, [vp, P(u64)]),
    "hv_vg_create": (i32, [vp, u32, P(vp)]),
    "hv_vg_destroy": (None, [vp]),
    "hv_vg_prepare": (i32, [vp, P(VgFrameData), vp, u32, u32, f32, vp, u32]),
    "hv_vg_execute": (i32, [vp, vp]),
    "hv_vg_execute_timed": (i32, [vp, vp, P(f32)]),
    "hv_vg_outputs": (i32, [vp, P(VgOutputs)]),
    "hv_vg_read": (i32, [vp, vp, vp, u32, vp, vp, u32]),
    "hv_vg_build_work_items": (i32, [vp, u32, vp, u32, P(u32)]),
    "hv_vg_scene_create": (i32, [vp, P(VgFrameData), P(vp)]),
    "hv_vg_scene_destroy": (None, [vp]),
    "hv_indirect_dispatch_create": (i32, [vp, P(vp)]),
    "hv_indirect_dispatch_destroy": (None, [vp]),
    "hv_indirect_dispatch_prepare": (i32, [vp, vp, u32]),
    "hv_indirect_dispatch_uniforms": (i32, [vp, vp]),
    "hv_indirect_dispatch_execute": (i32, [vp, P(DrawBuffers), vp]),
    "hv_occlusion_cull_create": (i32, [vp, u32, u32, P(vp)]),
    "hv_occlusion_cull_destroy": (None, [vp]),
    "hv_occlusion_cull_prepare": (i32, [vp, vp, u32, u64]),
    "hv_occlusion_cull_params": (i32, [vp, vp]),
    "hv_occlusion_cull_execute": (i32, [vp, P(DrawBuffers), vp, vp]),
    "hv_compact_draws": (i32, [vp, vp, u32, vp, vp, vp]),
    "hv_scene_rebuild_create": (i32, [vp, P(vp)]),
    "hv_scene_rebuild_destroy": (None, [vp]),
    "hv_scene_rebuild_execute": (i32, [vp, P(RebuildInputs), P(RebuildOutputs), u32, vp]),
    "hv_shadow_matrix_create": (i32, [vp, u32, P(vp)]),
    "hv_shadow_matrix_destroy": (None, [vp]),
    "hv_shadow_matrix_execute": (i32, [vp, P(ShadowMatrixBuffers), vp]),
    "hv_shadow_cull_create": (i32, [vp, u32, u32, P(vp)]),
    "hv_shadow_cull_destroy": (None, [vp]),
    "hv_shadow_cull_execute": (i32, [vp, P(ShadowCullBuffers), vp]),
    "hv_shadow_cull_outputs": (i32, [vp, P(vp), P(vp)]),
    "hv_shadow_cull_read": (i32, [vp, vp, vp]),
    "hv_server_create": (i32, [vp, P(ServerConfig), P(vp)]),
    "hv_server_destroy": (None, [vp]),
    "hv_server_submit": (i32, [vp, P(ServerRequest), P(u64)]),
    "hv_server_wait": (i32, [vp, u64, i32, P(u32)]),
    "hv_server_stats_get": (i32, [vp, P(ServerStats)]),
    "hv_server_stats_reset": (i32, [vp]),
    "hv_server_run_trace": (i32, [vp, P(ServerRequest), P(C.c_double), u32, P(C.c_double), P(C.c_double)]),
    "hv_ctx_set_shapes": (i32, [vp, u32, u32, u32]),
    "hv_ctx_get_shapes": (i32, [vp, P(u32), P(u32), P(u32)]),
    "hv_autotune_batch": (i32, [vp, u32, P(TuneResult), u32, P(u32), P(TuneResult)]),
    "hv_autotune_batch_replay": (i32, [vp, u32, P(TuneResult), u32, P(u32), P(TuneResult)]),
    "hv_autotune_hiz": (i32, [vp, u32, P(f32), P(u32)]),
}

_lock = threading.Lock()
_lib: C.CDLL | None = None


class HvError(RuntimeError):
    """A failing hv_* call: carries the status code and hv_last_error()."""

    def __init__(self, fn: str, status: int, message: str):
        super().__init__(f"{fn} -> {status}: {message}")
        self.status = status


def declared_symbols() -> list[str]:
    return sorted(_SIGS)


def load() -> C.CDLL:
    """Load libhv.so (once). Raises if it is missing or lacks a symbol."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m helio_amd._build` "
                "(no CPU fallback exists for the visibility path)")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)  # AttributeError if the symbol is absent
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def call(name: str, *args) -> int:
    """Call an hv_* entry point that returns hv_status; raise on failure."""
    lib = load()
    st = getattr(lib, name)(*args)
    if st != 0:
        msg = lib.hv_last_error().decode(errors="replace")
        raise HvError(name, st, msg)
    return st


def ptr(a) -> int | None:
    """Host pointer of a numpy array (C-contiguous) or None."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data
