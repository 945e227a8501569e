"""numpy mirrors of the GPU-shared layouts (libhelio) used at the C ABI.

Each dtype has the exact byte size and field offsets of the reference struct
it mirrors; tests compare them against the reference's own layout tests.
"""
from __future__ import annotations

import numpy as np

# GpuCameraUniforms, crates/libhelio/src/camera.rs:12-29 (368 bytes).
CAMERA = np.dtype([
    ("view", "<f4", 16),
    ("proj", "<f4", 16),
    ("view_proj", "<f4", 16),
    ("inv_view_proj", "<f4", 16),
    ("position_near", "<f4", 4),
    ("forward_far", "<f4", 4),
    ("jitter_frame", "<f4", 4),
    ("prev_view_proj", "<f4", 16),
])

# GpuInstanceData, crates/libhelio/src/instance.rs:30-45 (144 bytes).
INSTANCE = np.dtype([
    ("model", "<f4", 16),
    ("normal_mat", "<f4", 12),
    ("bounds", "<f4", 4),
    ("mesh_id", "<u4"),
    ("material_id", "<u4"),
    ("flags", "<u4"),
    ("lightmap_index", "<u4"),
])

# GpuInstanceAabb, crates/libhelio/src/instance.rs:60-65 (32 bytes).
AABB = np.dtype([("min", "<f4", 3), ("_pad0", "<f4"), ("max", "<f4", 3), ("_pad1", "<f4")])

# GpuDrawCall, crates/libhelio/src/draw.rs:28-36 (20 bytes).
DRAW_CALL = np.dtype([
    ("index_count", "<u4"),
    ("first_index", "<u4"),
    ("vertex_offset", "<i4"),
    ("first_instance", "<u4"),
    ("instance_count", "<u4"),
])

# DrawIndexedIndirectArgs, crates/libhelio/src/draw.rs:44-50 (20 bytes).
DRAW_INDEXED_INDIRECT = np.dtype([
    ("index_count", "<u4"),
    ("instance_count", "<u4"),
    ("first_index", "<u4"),
    ("base_vertex", "<i4"),
    ("first_instance", "<u4"),
])

# GpuMeshletEntry, crates/libhelio/src/meshlet.rs:43-72 (64 bytes).
MESHLET = np.dtype([
    ("center", "<f4", 3),
    ("radius", "<f4"),
    ("cone_apex", "<f4", 3),
    ("cone_cutoff", "<f4"),
    ("cone_axis", "<f4", 3),
    ("lod_error", "<f4"),
    ("first_index", "<u4"),
    ("index_count", "<u4"),
    ("vertex_offset", "<i4"),
    ("instance_index", "<u4"),
])

# GpuVgObject, crates/libhelio/src/meshlet.rs:81-102 (128 bytes).
VG_OBJECT = np.dtype([
    ("instance_index", "<u4"),
    ("lod_count", "<u4"),
    ("max_meshlet_count", "<u4"),
    ("selected_lod_plus_one", "<u4"),
    ("local_bounds", "<f4", 4),
    ("lod_errors", "<f4", 8),
    ("lod_first_meshlets", "<u4", 8),
    ("lod_meshlet_counts", "<u4", 8),
])

# GpuVgDraw, crates/libhelio/src/meshlet.rs:111-116 (16 bytes).
VG_DRAW = np.dtype([
    ("instance_index", "<u4"),
    ("meshlet_index", "<u4"),
    ("lod_level", "<u4"),
    ("reserved", "<u4"),
])

# GpuVgWorkItem, crates/libhelio/src/meshlet.rs:126-129 (8 bytes).
VG_WORK_ITEM = np.dtype([("object_index", "<u4"), ("local_meshlet_base", "<u4")])

# InstanceCullData, crates/helio-pass-virtual-geometry/src/lib.rs:122-129 (16 bytes).
INSTANCE_CULL = np.dtype([
    ("max_scale", "<f4"),
    ("min_scale", "<f4"),
    ("cone_cull_enabled", "<u4"),
    ("valid_transform", "<u4"),
])

# CullUniforms (indirect dispatch), crates/helio-pass-indirect-dispatch/src/lib.rs:16-22 (112 bytes).
ID_CULL_UNIFORMS = np.dtype([("frustum_planes", "<f4", (6, 4)), ("draw_count", "<u4"), ("_pad", "<u4", 3)])

# CullParams (occlusion), crates/helio-pass-occlusion-cull/src/lib.rs:20-35 (56 bytes).
OCCL_CULL_PARAMS = np.dtype([
    ("screen_width", "<u4"),
    ("screen_height", "<u4"),
    ("draw_count", "<u4"),
    ("hiz_mip_count", "<u4"),
    ("static_hiz_available", "<u4"),
    ("grid_resolution", "<u4", 3),
    ("world_bounds_min", "<f4", 3),
    ("world_bounds_max", "<f4", 3),
])

# CullUniforms (virtual geometry), crates/helio-pass-virtual-geometry/src/lib.rs:257-272 (48 bytes).
VG_CULL_UNIFORMS = np.dtype([
    ("object_count", "<u4"),
    ("screen_width", "<u4"),
    ("screen_height", "<u4"),
    ("hiz_mip_count", "<u4"),
    ("draw_capacity", "<u4"),
    ("lod_error_threshold_px", "<f4"),
    ("object_dispatch_width", "<u4"),
    ("work_item_count", "<u4"),
    ("work_dispatch_width", "<u4"),
    ("_pad", "<u4", 3),
])

# HiZUniforms, crates/helio-pass-hiz/src/lib.rs:29-34 (16 bytes).
HIZ_UNIFORMS = np.dtype([("src_size", "<u4", 2), ("dst_size", "<u4", 2)])

# hv_mesh_lod: (index_count, first_index, base_vertex) of one (mesh, LOD) bucket.
MESH_LOD = np.dtype([("index_count", "<u4"), ("first_index", "<u4"), ("base_vertex", "<i4")])

# GpuLight, crates/libhelio/src/light.rs:29-45 (64 bytes).
LIGHT = np.dtype([
    ("position_range", "<f4", 4),
    ("direction_outer", "<f4", 4),
    ("color_intensity", "<f4", 4),
    ("shadow_index", "<u4"),
    ("light_type", "<u4"),
    ("inner_angle", "<f4"),
    ("_pad", "<u4"),
])
LIGHT_DIRECTIONAL, LIGHT_POINT, LIGHT_SPOT = 0, 1, 2
SHADOW_NO_SHADOW = 0xFFFFFFFF
SHADOW_MAX_FACES = 256
SHADOW_MAX_DRAWS_PER_FACE = 4096
SHADOW_PLANES_REFERENCE = 0  # shadow_cull.wgsl's column vectors
SHADOW_PLANES_ROWS = 1       # Gribb-Hartmann rows

STAT_NAMES = (
    "total", "frustum_culled", "subpixel_culled", "frustum_visible",
    "occlusion_culled", "shadow_total", "shadow_frustum_visible", "shadow_occlusion_culled",
)

HIZ_MAX = 0
HIZ_MIN = 1
CULL_OCCLUSION = 1
CULL_REVERSED_Z = 2
VG_CLEAR_INDIRECT = 0x100

# Instance/draw rebuild (crates/helio/src/scene/objects/rebuild.rs).
# hv_object_info: the ObjectRecord fields the rebuild reads besides instance,
# aabb and draw (crates/helio/src/scene/types.rs:157-187). 16 bytes.
OBJECT_INFO = np.dtype([("groups", "<u8"), ("material_slot", "<u4"), ("movability", "<u4")])
# (graph_hash, material_class) of one material. 16 bytes.
MATERIAL_KEY = np.dtype([("graph_hash", "<u8"), ("material_class", "<u4"), ("_pad", "<u4")])
# GpuScene::material_class_ranges entry (crates/helio-core/src/scene/gpu_scene.rs:236-237). 24 bytes.
MATERIAL_RANGE = np.dtype([("graph_hash", "<u8"), ("material_class", "<u4"), ("start", "<u4"), ("count", "<u4"),
                           ("_pad", "<u4")])
MOVABILITY_STATIC, MOVABILITY_STATIONARY, MOVABILITY_MOVABLE = 0, 1, 2
REBUILD_PERSISTENT, REBUILD_OPTIMIZED = 0, 1
REBUILD_MAX_OBJECTS = 1 << 28

