/*
 * hv — MI355X-native GPU-driven visibility: instance frustum culling with
 * distance LOD, the Hi-Z pyramid, Hi-Z occlusion, meshlet (virtual-geometry)
 * culling and stream compaction into DrawIndexedIndirect argument buffers.
 *
 * This is the C ABI of libhv.so. Every struct below keeps the byte layout of the
 * GPU-shared struct it mirrors in the reference (t4ce/Helio, crate libhelio and
 * the pass crates); every function cites the reference interface it replaces.
 * Only plain pointers, sizes and status codes cross this boundary.
 *
 * Conventions
 *  - Every entry point returns hv_status (HV_OK == 0). On failure the message is
 *    available from hv_last_error() (thread-local). No exception crosses the ABI.
 *  - `stream` arguments are hipStream_t values passed as void*; NULL means the
 *    context's own stream.
 *  - `*_prepare` never touches the device queue except for small uniform
 *    uploads; `*_execute` only enqueues work (record-only, no host blocking),
 *    mirroring RenderPass::prepare / RenderPass::execute.
 *  - Matrices are column-major float[16] (glam::Mat4::to_cols_array()).
 *  - Handles are not internally synchronised: one caller thread per handle.
 *    hv_server is the exception and is safe to call from any thread.
 */
#ifndef HV_HV_H
#define HV_HV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Status codes. Replaces helio_core::Result / Error                          */
/* (crates/helio-core/src/error.rs:37,62) and the assert!/expect panics of    */
/* the passes (crates/helio-pass-virtual-geometry/src/rendering.rs:836-845).  */
/* ------------------------------------------------------------------------- */
typedef int32_t hv_status;
#define HV_OK 0
#define HV_EINVAL (-1)   /* bad argument, size or layout */
#define HV_ENOMEM (-2)   /* device allocation failed or admission refused */
#define HV_ELIMIT (-3)   /* grid, capacity or table limit exceeded */
#define HV_EDEVICE (-4)  /* HIP runtime error */
#define HV_EAGAIN (-5)   /* result not ready yet (server poll) */
#define HV_ESTATE (-6)   /* call order violated (execute before prepare, ...) */

/* ------------------------------------------------------------------------- */
/* GPU-shared layouts (libhelio). Sizes are static-asserted in the library.   */
/* ------------------------------------------------------------------------- */

/* GpuCameraUniforms, crates/libhelio/src/camera.rs:12-29. 368 bytes (the
 * reference doc comment says 256; the struct is 4 mat4 + 3 vec4 + 1 mat4). */
typedef struct hv_camera_uniforms {
    float view[16];           /* @0   */
    float proj[16];           /* @64  */
    float view_proj[16];      /* @128 */
    float inv_view_proj[16];  /* @192 */
    float position_near[4];   /* @256 xyz = position, w = near */
    float forward_far[4];     /* @272 xyz = forward,  w = far  */
    float jitter_frame[4];    /* @288 */
    float prev_view_proj[16]; /* @304 */
} hv_camera_uniforms;

/* GpuInstanceData, crates/libhelio/src/instance.rs:30-45. 144 bytes. */
typedef struct hv_instance_data {
    float model[16];      /* @0   column-major model matrix */
    float normal_mat[12]; /* @64  3 x vec4 (xyz used) */
    float bounds[4];      /* @112 world-space sphere: xyz center, w radius */
    uint32_t mesh_id;     /* @128 */
    uint32_t material_id; /* @132 */
    uint32_t flags;       /* @136 bit0 = casts shadow */
    uint32_t lightmap_index; /* @140 */
} hv_instance_data;

/* GpuInstanceAabb, crates/libhelio/src/instance.rs:60-65. 32 bytes. */
typedef struct hv_instance_aabb {
    float min[3];
    float _pad0;
    float max[3];
    float _pad1;
} hv_instance_aabb;

/* GpuDrawCall, crates/libhelio/src/draw.rs:28-36. 20 bytes. */
typedef struct hv_draw_call {
    uint32_t index_count;
    uint32_t first_index;
    int32_t vertex_offset;
    uint32_t first_instance;
    uint32_t instance_count;
} hv_draw_call;

/* DrawIndexedIndirectArgs, crates/libhelio/src/draw.rs:44-50. 20 bytes. */
typedef struct hv_draw_indexed_indirect {
    uint32_t index_count;
    uint32_t instance_count;
    uint32_t first_index;
    int32_t base_vertex;
    uint32_t first_instance;
} hv_draw_indexed_indirect;

/* GpuMeshletEntry, crates/libhelio/src/meshlet.rs:43-72. 64 bytes. */
typedef struct hv_meshlet_entry {
    float center[3];
    float radius;
    float cone_apex[3];
    float cone_cutoff; /* > 1 disables the cone test */
    float cone_axis[3];
    float lod_error;   /* not read by the cull (vg_cull.wgsl:135-264) */
    uint32_t first_index;
    uint32_t index_count;
    int32_t vertex_offset;
    uint32_t instance_index; /* reserved */
} hv_meshlet_entry;

#define HV_VG_LOD_LEVELS 8
#define HV_VG_MESHLETS_PER_WORK_ITEM 64
#define HV_VG_COUNTER_COUNT 11

/* GpuVgObject, crates/libhelio/src/meshlet.rs:81-102. 128 bytes. */
typedef struct hv_vg_object {
    uint32_t instance_index;
    uint32_t lod_count;
    uint32_t max_meshlet_count;
    uint32_t selected_lod_plus_one; /* GPU scratch (reference field `reserved`) */
    float local_bounds[4];
    float lod_errors[HV_VG_LOD_LEVELS];
    uint32_t lod_first_meshlets[HV_VG_LOD_LEVELS];
    uint32_t lod_meshlet_counts[HV_VG_LOD_LEVELS];
} hv_vg_object;

/* GpuVgDraw, crates/libhelio/src/meshlet.rs:111-116. 16 bytes. */
typedef struct hv_vg_draw {
    uint32_t instance_index;
    uint32_t meshlet_index;
    uint32_t lod_level;
    uint32_t reserved;
} hv_vg_draw;

/* GpuVgWorkItem, crates/libhelio/src/meshlet.rs:126-129. 8 bytes. */
typedef struct hv_vg_work_item {
    uint32_t object_index;
    uint32_t local_meshlet_base;
} hv_vg_work_item;

/* InstanceCullData, crates/helio-pass-virtual-geometry/src/lib.rs:122-129. 16 bytes. */
typedef struct hv_instance_cull_data {
    float max_scale;
    float min_scale;
    uint32_t cone_cull_enabled;
    uint32_t valid_transform;
} hv_instance_cull_data;

/* CullUniforms of the indirect-dispatch pass,
 * crates/helio-pass-indirect-dispatch/src/lib.rs:16-22. 112 bytes. */
typedef struct hv_id_cull_uniforms {
    float frustum_planes[6][4];
    uint32_t draw_count;
    uint32_t _pad[3];
} hv_id_cull_uniforms;

/* CullParams of the occlusion pass,
 * crates/helio-pass-occlusion-cull/src/lib.rs:20-35. 56 bytes. */
typedef struct hv_occl_cull_params {
    uint32_t screen_width;
    uint32_t screen_height;
    uint32_t draw_count;
    uint32_t hiz_mip_count;
    uint32_t static_hiz_available; /* must be 0: the baked-PVS branch is not built */
    uint32_t grid_resolution[3];
    float world_bounds_min[3];
    float world_bounds_max[3];
} hv_occl_cull_params;

/* CullUniforms of the virtual-geometry pass,
 * crates/helio-pass-virtual-geometry/src/lib.rs:257-272. 48 bytes. */
typedef struct hv_vg_cull_uniforms {
    uint32_t object_count;
    uint32_t screen_width;
    uint32_t screen_height;
    uint32_t hiz_mip_count;
    uint32_t draw_capacity;
    float lod_error_threshold_px;
    uint32_t object_dispatch_width;
    uint32_t work_item_count;
    uint32_t work_dispatch_width;
    uint32_t _pad[3];
} hv_vg_cull_uniforms;

/* HiZUniforms, crates/helio-pass-hiz/src/lib.rs:29-34. 16 bytes. */
typedef struct hv_hiz_uniforms {
    uint32_t src_size[2];
    uint32_t dst_size[2];
} hv_hiz_uniforms;

/* Cull statistics, 8 x u32, layout of indirect_dispatch.wgsl:69-77:
 * 0 total, 1 frustum_culled, 2 subpixel_culled, 3 frustum_visible,
 * 4 occlusion_culled, 5 shadow_total, 6 shadow_frustum_visible,
 * 7 shadow_occlusion_culled. */
#define HV_STAT_COUNT 8

/* ------------------------------------------------------------------------- */
/* Per-(mesh, LOD) draw range used to fill bucket args. No reference struct:  */
/* the compacting per-instance cull is this library's extension of           */
/* indirect_dispatch.wgsl (which culls per draw call and never selects LOD). */
/* ------------------------------------------------------------------------- */
typedef struct hv_mesh_lod {
    uint32_t index_count;
    uint32_t first_index;
    int32_t base_vertex;
} hv_mesh_lod;

/* ------------------------------------------------------------------------- */
/* Context, errors, device memory plumbing                                   */
/* ------------------------------------------------------------------------- */
typedef struct hv_ctx hv_ctx;

/* Thread-local message of the last failure on this thread. Never NULL. */
const char* hv_last_error(void);

/* Device context: owns a stream, the launch-shape table and the device
 * properties. Replaces the wgpu::Device/Queue pair handed to every pass
 * (crates/helio-core/src/context.rs:450-494). */
hv_status hv_ctx_create(int32_t device, hv_ctx** out);
void hv_ctx_destroy(hv_ctx* ctx);
hv_status hv_ctx_stream(hv_ctx* ctx, void** stream);
hv_status hv_ctx_synchronize(hv_ctx* ctx);
/* Extra non-blocking streams at the default priority, for callers that keep
 * several batches in flight (the encoder/queue split of the reference has no
 * counterpart: wgpu submits in order). */
hv_status hv_stream_create(hv_ctx* ctx, void** stream);
hv_status hv_stream_destroy(hv_ctx* ctx, void* stream);
hv_status hv_stream_synchronize(void* stream);
/* Number of compute units and the L2/LLC sizes seen by the library. */
hv_status hv_ctx_device_info(hv_ctx* ctx, uint32_t* cu_count, uint64_t* hbm_bytes);

/* Device buffers, the counterpart of wgpu::Device::create_buffer and
 * Queue::write_buffer (crates/helio-core/src/context.rs:498). */
hv_status hv_device_alloc(hv_ctx* ctx, uint64_t bytes, void** dptr);
hv_status hv_device_free(hv_ctx* ctx, void* dptr);
hv_status hv_memcpy_h2d(hv_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream);
hv_status hv_memcpy_d2h(hv_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream);
hv_status hv_memset_d32(hv_ctx* ctx, void* dst, uint32_t value, uint64_t count, void* stream);
/* Device-to-device copy (CommandEncoder::copy_buffer_to_buffer, as used at
 * crates/helio-pass-virtual-geometry/src/rendering.rs:1168); synchronises. */
hv_status hv_memcpy_d2d(hv_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream);

/* Stream-ordered timing: record a HIP event on `stream` into slot `slot`
 * (0..63) and read the elapsed ms between two recorded slots. */
hv_status hv_timer_record(hv_ctx* ctx, uint32_t slot, void* stream);
hv_status hv_timer_elapsed_ms(hv_ctx* ctx, uint32_t slot_begin, uint32_t slot_end, float* ms);

/* ------------------------------------------------------------------------- */
/* Math helpers shared with the host side of the reference passes            */
/* ------------------------------------------------------------------------- */

/* extract_frustum_planes, crates/helio-pass-indirect-dispatch/src/lib.rs:271-300:
 * Gribb-Hartmann rows of view_proj, normalised when |n| > 1e-10.
 * Order L, R, B, T, near (row 2, z in [0,1]), far. */
hv_status hv_extract_frustum_planes(const float view_proj[16], float planes[24]);

/* mip_levels, crates/helio-pass-hiz/src/lib.rs:376-379 and
 * crates/helio-pass-occlusion-cull/src/lib.rs:413-416. */
uint32_t hv_mip_levels(uint32_t width, uint32_t height);

/* InstanceCullData::from_instance,
 * crates/helio-pass-virtual-geometry/src/lib.rs:131-168 (Rust f32 semantics). */
hv_status hv_instance_cull_data_from_instances(const hv_instance_data* instances, uint32_t count,
                                               hv_instance_cull_data* out);

/* ------------------------------------------------------------------------- */
/* Resident scene (GpuScene / GrowableBuffer residency,                     */
/* crates/helio-core/src/scene/gpu_scene.rs:142-248,                         */
/* crates/helio-core/src/scene/managers.rs:19-249).                          */
/* Instances are converted AoS -> SoA on the device at upload: bounds as     */
/* float4 (16 B) and a 4-byte meta word (mesh_id | casts_shadow << 31), so   */
/* the cull reads 20 B per instance instead of 144 + 32 B.                   */
/* A pass or batch prepared on a scene keeps its device pointers: an upload  */
/* that grows the scene, switches it to explicit AABBs, a new LOD table or a */
/* rebuild into it makes the next execute return HV_ESTATE until prepared    */
/* again (the reference rebuilds its bind groups when a buffer changes,      */
/* crates/helio-pass-indirect-dispatch/src/lib.rs:200-244). Uploads within   */
/* the capacity need no new prepare. Likewise a pyramid rebound to other    */
/* depth (hv_hiz_bind_depth) after an occlusion or VG prepare.                */
/* ------------------------------------------------------------------------- */
typedef struct hv_scene hv_scene;

hv_status hv_scene_create(hv_ctx* ctx, uint32_t capacity, hv_scene** out);
void hv_scene_destroy(hv_scene* scene);
/* Upload instances [first, first + count) (a dirty range; the scene grows x2
 * like GrowableBuffer). `aabbs` may be NULL, meaning sphere_to_aabb(bounds)
 * (crates/helio/src/scene/helpers.rs:54-65); a non-NULL array that differs from
 * sphere_to_aabb anywhere switches the scene to explicit AABBs. mesh_id must be
 * below 2^31. */
hv_status hv_scene_upload(hv_scene* scene, const hv_instance_data* instances,
                          const hv_instance_aabb* aabbs, uint32_t first, uint32_t count,
                          void* stream);
/* Upload already-SoA data: bounds (count x float4), mesh ids and flags. */
hv_status hv_scene_upload_soa(hv_scene* scene, const float* bounds4, const uint32_t* mesh_ids,
                              const uint32_t* flags, uint32_t first, uint32_t count, void* stream);
/* Upload instances [first, first + count) from a DEVICE array of
 * GpuInstanceData (the reference's GpuScene instance buffer, no PCIe copy):
 * converted to the resident SoA on the device; sphere-derived AABBs only.
 * Synchronises once to validate the mesh ids (< 2^31). */
hv_status hv_scene_upload_device(hv_scene* scene, const hv_instance_data* instances, uint32_t first,
                                 uint32_t count, void* stream);
/* GrowableBuffer::swap_remove (crates/helio-core/src/scene/managers.rs:147-160):
 * the last instance moves into `index`, the count drops by one. */
hv_status hv_scene_swap_remove(hv_scene* scene, uint32_t index, void* stream);
/* GrowableBuffer::truncate (managers.rs:162-180): keeps the first new_len instances. */
hv_status hv_scene_truncate(hv_scene* scene, uint32_t new_len);
/* Blocking readback of the resident SoA (bounds: count x float4; meta:
 * mesh_id | casts_shadow << 31); either pointer may be NULL. */
hv_status hv_scene_read_soa(hv_scene* scene, float* bounds4, uint32_t* meta, uint32_t first, uint32_t count);
/* Per-(mesh, LOD) draw ranges: table[mesh * lod_count + lod]. */
hv_status hv_scene_set_mesh_lods(hv_scene* scene, const hv_mesh_lod* table, uint32_t mesh_count,
                                 uint32_t lod_count);
hv_status hv_scene_info(const hv_scene* scene, uint32_t* instance_count, uint32_t* mesh_count,
                        uint32_t* lod_count, uint64_t* device_bytes);

/* ------------------------------------------------------------------------- */
/* Instance/draw rebuild (Scene::rebuild_instance_buffers_persistent and      */
/* _optimized, crates/helio/src/scene/objects/rebuild.rs:91-191,234-361, and  */
/* rebuild_shadow_partition_buffers, rebuild.rs:374-425), on the device.      */
/* ------------------------------------------------------------------------- */

/* The fields of one ObjectRecord (crates/helio/src/scene/types.rs:157-187)
 * that the rebuild reads besides instance, aabb and draw. 16 bytes. */
typedef struct hv_object_info {
    uint64_t groups;          /* GroupMask (crates/helio/src/groups.rs:99) */
    uint32_t material_slot;   /* row of the material table; >= material_count reads (class 0, graph_hash 0),
                                 as the reference's failed materials.get() */
    uint32_t movability;      /* Movability (crates/libhelio/src/movability.rs): 0 static, 1 stationary,
                                 2 movable; only 2 goes to the movable shadow list */
} hv_object_info;

/* (material_class, graph_hash) of one material, the pair rebuild.rs:108-117 sorts by. 16 bytes. */
typedef struct hv_material_key {
    uint64_t graph_hash;
    uint32_t material_class;
    uint32_t _pad;
} hv_material_key;

/* One entry of GpuScene::material_class_ranges (class, graph_hash, start, count),
 * crates/helio-core/src/scene/gpu_scene.rs:236-237. 24 bytes. */
typedef struct hv_material_range {
    uint64_t graph_hash;
    uint32_t material_class;
    uint32_t start;
    uint32_t count;
    uint32_t _pad;
} hv_material_range;

/* Device inputs, in the dense object order. */
typedef struct hv_rebuild_inputs {
    const hv_instance_data* instances;   /* 16-byte aligned */
    const hv_instance_aabb* aabbs;       /* 16-byte aligned */
    const hv_draw_call* draws;           /* the draw templates (first_instance, instance_count unused) */
    const hv_object_info* objects;       /* 8-byte aligned */
    const hv_material_key* materials;    /* 8-byte aligned; may be NULL when material_count is 0 */
    uint32_t object_count;
    uint32_t material_count;
    uint64_t group_hidden;               /* Scene::group_hidden */
} hv_rebuild_inputs;

/* Device outputs; each array holds object_count entries (the worst case). */
typedef struct hv_rebuild_outputs {
    hv_instance_data* instances;         /* sorted instances (16-byte aligned) */
    hv_instance_aabb* aabbs;             /* sorted aabbs (16-byte aligned) */
    hv_draw_call* draw_calls;            /* one per draw group */
    hv_draw_indexed_indirect* indirect;  /* one per draw group */
    uint32_t* visibility;                /* per sorted slot: object_is_visible(groups, group_hidden) */
    uint32_t* gpu_slots;                 /* per dense object: its sorted slot (ObjectRecord::gpu_slot) */
    hv_material_range* ranges;           /* material_class_ranges (8-byte aligned) */
    hv_draw_indexed_indirect* shadow_static;   /* static + stationary objects, dense order */
    hv_draw_indexed_indirect* shadow_movable;  /* movable objects, dense order */
    uint32_t* counts;                    /* [0] draw groups, [1] ranges, [2] shadow_static_draw_count,
                                            [3] shadow_movable_draw_count */
    hv_scene* scene;                     /* optional: resident SoA replaced by the sorted instances
                                            (count = object_count, sphere-derived AABBs) */
} hv_rebuild_outputs;

#define HV_REBUILD_PERSISTENT 0u  /* sort by (class, graph_hash); one draw per object */
#define HV_REBUILD_OPTIMIZED 1u   /* sort by (class, graph_hash, mesh_id, material_id); one draw per run of
                                     equal (mesh_id, material_id) with instance_count = run length */
#define HV_REBUILD_MAX_OBJECTS (1u << 28)
#define HV_REBUILD_MAX_MATERIALS (1u << 24)

typedef struct hv_scene_rebuild hv_scene_rebuild;

hv_status hv_scene_rebuild_create(hv_ctx* ctx, hv_scene_rebuild** out);
void hv_scene_rebuild_destroy(hv_scene_rebuild* rb);
/* Both sorts are stable (ties keep the dense order, as Rust's sort_by_key).
 * Ranges run over objects (persistent) or over draw groups (optimized) and
 * split where (class, graph_hash) changes between neighbours; a group takes
 * the class, graph_hash and draw template of its first object. With zero
 * objects counts[0..1] become 0 and counts[2..3] are left as they are (the
 * reference returns before rebuilding the shadow partition). Synchronises
 * the stream once (key widths come from the material ranks and the largest
 * mesh and material ids); the rest is enqueued. No output may overlap an
 * input (HV_EINVAL). */
hv_status hv_scene_rebuild_execute(hv_scene_rebuild* rb, const hv_rebuild_inputs* in, const hv_rebuild_outputs* out,
                                   uint32_t mode, void* stream);

/* ------------------------------------------------------------------------- */
/* Hi-Z pyramid (HiZBuildPass, crates/helio-pass-hiz/src/lib.rs:45-548;       */
/* hiz_depth_copy.wgsl:13-19, hiz_build.wgsl:12-23).                         */
/* Level 0 is the depth buffer itself (no copy); levels 1.. are packed in one */
/* device buffer. HV_HIZ_MAX is the reference (standard Z, max reduction);   */
/* HV_HIZ_MIN is its mirror for reversed Z (min reduction).                   */
/* ------------------------------------------------------------------------- */
#define HV_HIZ_MAX 0u
#define HV_HIZ_MIN 1u

typedef struct hv_hiz hv_hiz;

hv_status hv_hiz_create(hv_ctx* ctx, uint32_t width, uint32_t height, uint32_t mode, hv_hiz** out);
void hv_hiz_destroy(hv_hiz* hiz);
/* Bind a caller-owned device depth buffer (width x height f32, row-major) as
 * level 0, or pass NULL to use the pyramid's own level-0 buffer. */
hv_status hv_hiz_bind_depth(hv_hiz* hiz, const float* depth_device);
/* Copy host depth into the pyramid's own level-0 buffer. */
hv_status hv_hiz_upload_depth(hv_hiz* hiz, const float* depth_host, void* stream);
/* Camera-static reuse (crates/helio-pass-hiz/src/lib.rs:472-478): when
 * `camera_generation` equals the one of the last build, execute is a no-op. */
hv_status hv_hiz_prepare(hv_hiz* hiz, uint64_t camera_generation);
hv_status hv_hiz_execute(hv_hiz* hiz, void* stream);
hv_status hv_hiz_levels(const hv_hiz* hiz, uint32_t* level_count);
hv_status hv_hiz_level_size(const hv_hiz* hiz, uint32_t level, uint32_t* width, uint32_t* height);
hv_status hv_hiz_read_level(hv_hiz* hiz, uint32_t level, float* out, uint64_t capacity_floats);

/* ------------------------------------------------------------------------- */
/* Per-instance cull + distance LOD + compaction (one view).                  */
/* Predicate: indirect_dispatch.wgsl:79-137 applied per instance; with        */
/* HV_CULL_OCCLUSION also occlusion_cull.wgsl:101-220 against an hv_hiz.      */
/* Output: visible instance ids grouped bucket-major (bucket = mesh_id *      */
/* lod_count + lod; order inside a bucket unspecified, like the reference's  */
/* atomic append), one hv_draw_indexed_indirect per bucket with               */
/* first_instance = exclusive prefix of bucket counts, and the 8 stats.       */
/* ------------------------------------------------------------------------- */
#define HV_CULL_OCCLUSION 1u  /* test survivors against the bound pyramid */
#define HV_CULL_REVERSED_Z 2u /* pyramid is HV_HIZ_MIN, compare mirrored */
#define HV_MAX_LOD_LEVELS 8

typedef struct hv_cull_config {
    uint32_t lod_count;                            /* 1..8 */
    float lod_distances[HV_MAX_LOD_LEVELS - 1];    /* ascending, lod = #{t : d >= t} */
    uint32_t flags;                                /* HV_CULL_* */
    uint32_t screen_width;                         /* used by occlusion */
    uint32_t screen_height;
} hv_cull_config;

typedef struct hv_inst_cull_pass hv_inst_cull_pass;

/* Device pointers of a view's outputs (valid until the next prepare). */
typedef struct hv_cull_outputs {
    uint32_t* ids;                   /* capacity: candidate count */
    hv_draw_indexed_indirect* args;  /* mesh_count * lod_count entries */
    uint32_t* stats;                 /* HV_STAT_COUNT entries, then visible count at [8] */
    uint32_t bucket_count;
    uint32_t id_capacity;
} hv_cull_outputs;

/* IndirectDispatchPass::new, crates/helio-pass-indirect-dispatch/src/lib.rs:39-161. */
hv_status hv_inst_cull_create(hv_ctx* ctx, const hv_cull_config* config, hv_inst_cull_pass** out);
void hv_inst_cull_destroy(hv_inst_cull_pass* pass);
/* IndirectDispatchPass::prepare, crates/helio-pass-indirect-dispatch/src/lib.rs:178-191:
 * frustum planes from camera->view_proj, candidate range [first, first+count). */
hv_status hv_inst_cull_prepare(hv_inst_cull_pass* pass, const hv_scene* scene,
                               const hv_camera_uniforms* camera, uint32_t first, uint32_t count,
                               const hv_hiz* hiz);
/* IndirectDispatchPass::execute, crates/helio-pass-indirect-dispatch/src/lib.rs:193-258. */
hv_status hv_inst_cull_execute(hv_inst_cull_pass* pass, void* stream);
hv_status hv_inst_cull_outputs(const hv_inst_cull_pass* pass, hv_cull_outputs* out);
/* Blocking readback (the async cull-stats readback of
 * crates/helio/src/renderer/render.rs:436-481 made synchronous). Any pointer
 * may be NULL to skip that output. */
hv_status hv_inst_cull_read(hv_inst_cull_pass* pass, uint32_t* ids, uint32_t ids_capacity,
                            uint32_t* id_count, hv_draw_indexed_indirect* args,
                            uint32_t args_capacity, uint32_t stats[HV_STAT_COUNT]);

/* One view of the cull above (hv_inst_cull_prepare + hv_inst_cull_execute in
 * one call) that allocates nothing: no hipMalloc, no hipFree, no host staging
 * copy. No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation). Same checks, kernels and results as the pass (bit-exact
 * ids per bucket, args, stats); the context's launch shape applies.
 * Inputs: a resident scene, the camera (read during the call), an optional
 * bound pyramid (HV_CULL_OCCLUSION). Outputs are the caller's device arrays:
 * ids (count entries), args (mesh_count * lod_count), counters (9 u32:
 * [0..7] stats, [8] visible ids). Only enqueues on `stream`.
 * Scratch: one static device array of HV_NOALLOC_SCRATCH_BYTES per GPU the
 * library runs on, held until the process ends, shared by every call on that
 * GPU: calls ordered on one stream are safe; calls that overlap on one GPU
 * (unordered streams) race. A call needs about 4 B per candidate + 4 B per
 * (row, bucket) with one row per classify workgroup (at most the resident
 * workgroups, e.g. 1,024), so 2^22 candidates at 1,024 buckets fit; a call
 * that does not fit returns HV_ELIMIT with the bytes it needs. */
#define HV_NOALLOC_SCRATCH_BYTES (32ull << 20)
hv_status hv_inst_cull_noalloc(hv_ctx* ctx, const hv_cull_config* config, const hv_scene* scene,
                               const hv_camera_uniforms* camera, uint32_t first, uint32_t count,
                               const hv_hiz* hiz, uint32_t* ids, hv_draw_indexed_indirect* args,
                               uint32_t* counters, void* stream);


/* ------------------------------------------------------------------------- */
/* Batched multi-view cull: many views (cameras, shadow cascades, cube faces) */
/* of possibly different resident scenes in ONE launch sequence, optionally  */
/* replayed from a captured hipGraph. Per-view results are identical to the  */
/* plain per-view pass. No reference counterpart: the reference dispatches   */
/* one compute pass per view (closest analog: shadow_cull.wgsl:57-96, which  */
/* loops over faces inside one dispatch).                                    */
/* ------------------------------------------------------------------------- */
typedef struct hv_view_request {
    const hv_scene* scene;
    const hv_camera_uniforms* camera;
    uint32_t first;   /* candidate range [first, first + count) of the scene */
    uint32_t count;
    const hv_hiz* hiz; /* required when the config has HV_CULL_OCCLUSION */
} hv_view_request;

typedef struct hv_batch hv_batch;

hv_status hv_batch_create(hv_ctx* ctx, const hv_cull_config* config, uint32_t max_views, hv_batch** out);
void hv_batch_destroy(hv_batch* batch);
/* Launch shape of the classify kernel: wg in {64,128,256,512}, ipt in {1,2,4,8}. */
hv_status hv_batch_set_shape(hv_batch* batch, uint32_t wg, uint32_t ipt);
hv_status hv_batch_prepare(hv_batch* batch, const hv_view_request* views, uint32_t count);
/* use_graph != 0: replay the captured graph (captured on first use). A replay
 * enqueues the graph only; the next hv_batch_wait / _read / _prepare records
 * completion on `stream`, which must stay valid until then. */
hv_status hv_batch_execute(hv_batch* batch, void* stream, uint32_t use_graph);
/* Eager execute with HIP events around the three kernels; synchronises and
 * returns ms[0] classify, ms[1] column scan (bucket offsets), ms[2] scatter. */
hv_status hv_batch_execute_timed(hv_batch* batch, void* stream, float ms[3]);
hv_status hv_batch_outputs(const hv_batch* batch, uint32_t view, hv_cull_outputs* out);
hv_status hv_batch_read(hv_batch* batch, uint32_t view, uint32_t* ids, uint32_t ids_capacity, uint32_t* id_count,
                        hv_draw_indexed_indirect* args, uint32_t args_capacity, uint32_t stats[HV_STAT_COUNT]);
/* Block until the last execute of this batch finished. */
hv_status hv_batch_wait(hv_batch* batch);
hv_status hv_batch_device_bytes(const hv_batch* batch, uint64_t* bytes);

/* The same timing hook for the plain per-view pass. */
hv_status hv_inst_cull_execute_timed(hv_inst_cull_pass* pass, void* stream, float ms[3]);
hv_status hv_inst_cull_set_shape(hv_inst_cull_pass* pass, uint32_t wg, uint32_t ipt);

/* ------------------------------------------------------------------------- */
/* Virtual geometry: per-object frustum + projected-error LOD select, then    */
/* per-meshlet frustum + normal cone + Hi-Z cull with atomic publication      */
/* (VirtualGeometryPass, crates/helio-pass-virtual-geometry/src/rendering.rs; */
/* vg_cull.wgsl:126-405).                                                     */
/* ------------------------------------------------------------------------- */

/* VgFrameData, crates/libhelio/src/frame.rs:491-517 (slices as typed pointers). */
typedef struct hv_vg_frame_data {
    const hv_meshlet_entry* meshlets;
    const hv_vg_object* objects;
    const hv_instance_data* instances;   /* one per object */
    const hv_vg_work_item* work_items;
    uint32_t meshlet_count;
    uint32_t object_count;
    uint32_t work_item_count;
    uint32_t max_draw_count;
    uint64_t buffer_version;
    uint64_t instance_version;
    uint32_t instance_dirty_start;
    uint32_t instance_dirty_count;
} hv_vg_frame_data;

typedef struct hv_vg_buffers {
    hv_draw_indexed_indirect* indirect; /* capacity entries; first_instance = slot */
    hv_vg_draw* metadata;               /* capacity entries */
    uint32_t* counters;                 /* HV_VG_COUNTER_COUNT: [0] attempted, [1] rejected,
                                           [2..9] selected-LOD histogram, [10] max LOD */
    uint32_t capacity;
} hv_vg_buffers;

typedef struct hv_vg_pass hv_vg_pass;

/* VirtualGeometryPass::new_with_budget (publication budget in meshlets,
 * crates/helio-pass-virtual-geometry/src/lib.rs:80-117). 0 is rejected like
 * VirtualGeometryBudget::new. */
hv_status hv_vg_create(hv_ctx* ctx, uint32_t publication_limit, hv_vg_pass** out);
void hv_vg_destroy(hv_vg_pass* pass);
/* VirtualGeometryPass::prepare (rendering.rs:708-955): uploads on a new
 * buffer_version, a dirty instance range on a new instance_version, derives
 * InstanceCullData on the host, writes the cull uniforms. flags: any of
 * HV_CULL_REVERSED_Z (the pyramid must then be HV_HIZ_MIN) and
 * HV_VG_CLEAR_INDIRECT: also zero the whole indirect buffer each execute, for
 * a consumer without multi_draw_indexed_indirect_count that draws every slot
 * (the !use_count_indirect path, rendering.rs:1127-1130,1203-1204). */
#define HV_VG_CLEAR_INDIRECT 0x100u
hv_status hv_vg_prepare(hv_vg_pass* pass, const hv_vg_frame_data* vg, const hv_camera_uniforms* camera,
                        uint32_t width, uint32_t height, float lod_error_threshold_px, const hv_hiz* hiz,
                        uint32_t flags);
/* VirtualGeometryPass::execute (rendering.rs:1031-1165): clear counters,
 * select objects, cull meshlets (selection and cull share one launch). */
hv_status hv_vg_execute(hv_vg_pass* pass, void* stream);
/* Execute with HIP events: ms[0] counter clear, ms[1] object select + meshlet
 * cull (one launch). Synchronises. */
hv_status hv_vg_execute_timed(hv_vg_pass* pass, void* stream, float ms[2]);
hv_status hv_vg_outputs(const hv_vg_pass* pass, hv_vg_buffers* out);
/* Blocking readback; any pointer may be NULL. selected: one u32 per object
 * (selected_lod_plus_one). */
hv_status hv_vg_read(hv_vg_pass* pass, hv_draw_indexed_indirect* indirect, hv_vg_draw* metadata,
                     uint32_t capacity, uint32_t counters[HV_VG_COUNTER_COUNT], uint32_t* selected,
                     uint32_t selected_capacity);
/* A device-resident virtual-geometry scene for the serving loop (its
 * meshlet scenes): the frame data uploaded once, InstanceCullData derived
 * at upload (InstanceCullData::from_instance, lib.rs:131-168). Requests on
 * one scene may run concurrently: they never write to it. */
typedef struct hv_vg_scene hv_vg_scene;
hv_status hv_vg_scene_create(hv_ctx* ctx, const hv_vg_frame_data* frame, hv_vg_scene** out);
void hv_vg_scene_destroy(hv_vg_scene* scene);

/* append_work_items for every object (crates/helio/src/scene/virtual_geometry/rebuild.rs:45-58):
 * one record per 64-meshlet span of max_meshlet_count. */
hv_status hv_vg_build_work_items(const hv_vg_object* objects, uint32_t object_count, hv_vg_work_item* out,
                                 uint32_t capacity, uint32_t* count);

/* ------------------------------------------------------------------------- */
/* Reference-exact per-draw passes over the reference's own buffers (AoS      */
/* GpuInstanceData, GpuDrawCall, GpuInstanceAabb, DrawIndexedIndirect), as    */
/* device pointers owned by the caller.                                       */
/* ------------------------------------------------------------------------- */
typedef struct hv_draw_buffers {
    const hv_instance_data* instances;  /* indexed by draw_calls[i].first_instance */
    const hv_instance_aabb* aabbs;      /* same indexing */
    const hv_draw_call* draw_calls;
    hv_draw_indexed_indirect* indirect; /* one per draw call */
    uint32_t* stats;                    /* HV_STAT_COUNT counters, cleared by the caller per frame */
    uint32_t draw_count;
    uint32_t instance_capacity;         /* entries in instances/aabbs; a first_instance beyond it
                                           reads zeros, like WGSL's robust buffer access */
} hv_draw_buffers;

typedef struct hv_indirect_dispatch_pass hv_indirect_dispatch_pass;

/* IndirectDispatchPass (crates/helio-pass-indirect-dispatch/src/lib.rs:24-258;
 * indirect_dispatch.wgsl:100-176): per draw call, the representative instance
 * (first_instance) is frustum (AABB unless degenerate, else sphere) and
 * sub-pixel tested; every draw gets its args, instance_count 0 when culled. */
hv_status hv_indirect_dispatch_create(hv_ctx* ctx, hv_indirect_dispatch_pass** out);
void hv_indirect_dispatch_destroy(hv_indirect_dispatch_pass* pass);
/* IndirectDispatchPass::prepare (:178-191): planes from the camera. */
hv_status hv_indirect_dispatch_prepare(hv_indirect_dispatch_pass* pass, const hv_camera_uniforms* camera,
                                       uint32_t draw_count);
/* The CullUniforms written by the last prepare (112 bytes). */
hv_status hv_indirect_dispatch_uniforms(const hv_indirect_dispatch_pass* pass, hv_id_cull_uniforms* out);
/* IndirectDispatchPass::execute (:193-258). */
hv_status hv_indirect_dispatch_execute(hv_indirect_dispatch_pass* pass, const hv_draw_buffers* buffers,
                                       void* stream);

typedef struct hv_occlusion_cull_pass hv_occlusion_cull_pass;

/* OcclusionCullPass (crates/helio-pass-occlusion-cull/src/lib.rs:37-411;
 * occlusion_cull.wgsl:146-220): per draw call, sphere -> screen rect -> mip ->
 * 4 nearest taps -> zero instance_count when occluded (stats 4 and 7 count
 * draws that were visible). The static-Hi-Z branch is not supported. */
hv_status hv_occlusion_cull_create(hv_ctx* ctx, uint32_t screen_width, uint32_t screen_height,
                                   hv_occlusion_cull_pass** out);
void hv_occlusion_cull_destroy(hv_occlusion_cull_pass* pass);
/* OcclusionCullPass::prepare (:297-316). frame_num 0 makes execute a no-op
 * (temporal Hi-Z has no pyramid yet, :319-323). */
hv_status hv_occlusion_cull_prepare(hv_occlusion_cull_pass* pass, const hv_camera_uniforms* camera,
                                    uint32_t draw_count, uint64_t frame_num);
hv_status hv_occlusion_cull_params(const hv_occlusion_cull_pass* pass, hv_occl_cull_params* out);
hv_status hv_occlusion_cull_execute(hv_occlusion_cull_pass* pass, const hv_draw_buffers* buffers, const hv_hiz* hiz,
                                    void* stream);

/* ------------------------------------------------------------------------- */
/* Indirect-args wire format. The per-instance pass writes one args record   */
/* per (mesh, LOD) bucket, empty buckets included (instance_count 0), which  */
/* multi_draw_indexed_indirect consumes as is. For                           */
/* multi_draw_indexed_indirect_count (the VG draw, rendering.rs:1194-1201,   */
/* and the shadow faces) this copies the non-empty draws, in order, to `out` */
/* and writes their number to count[0] (count at byte 0, as the VG draw      */
/* counter buffer keeps it, rendering.rs:648-661). Device pointers; out has  */
/* room for n records; n <= 2^20; out == args compacts in place (each draw   */
/* moves only towards the front, after its chunk was read). Enqueues on      */
/* `stream` only.                                                            */
/* ------------------------------------------------------------------------- */
hv_status hv_compact_draws(hv_ctx* ctx, const hv_draw_indexed_indirect* args, uint32_t n,
                           hv_draw_indexed_indirect* out, uint32_t* count, void* stream);

/* ------------------------------------------------------------------------- */
/* Multi-view producers: light-space view-projections for shadow faces       */
/* (ShadowMatrixPass) and per-face culling of the movable shadow draws       */
/* (ShadowCullPass), over device buffers owned by the caller.                */
/* ------------------------------------------------------------------------- */

/* GpuLight, crates/libhelio/src/light.rs:29-45. 64 bytes. */
typedef struct hv_light {
    float position_range[4];  /* @0  xyz position, w range */
    float direction_outer[4]; /* @16 xyz direction, w cos(outer angle) */
    float color_intensity[4]; /* @32 */
    uint32_t shadow_index;    /* @48 first of 6 shadow matrices; 0xFFFFFFFF = no shadow */
    uint32_t light_type;      /* @52 HV_LIGHT_* */
    float inner_angle;        /* @56 */
    uint32_t _pad;            /* @60 */
} hv_light;

#define HV_LIGHT_DIRECTIONAL 0u
#define HV_LIGHT_POINT 1u
#define HV_LIGHT_SPOT 2u
#define HV_SHADOW_NO_SHADOW 0xFFFFFFFFu
#define HV_SHADOW_FACES_PER_LIGHT 6u
#define HV_SHADOW_MAX_FACES 256u           /* MAX_FACES, shadow_cull.wgsl:5 */
#define HV_SHADOW_MAX_DRAWS_PER_FACE 4096u /* crates/helio-pass-shadow-cull/src/lib.rs:41 */

typedef struct hv_shadow_matrix_buffers {
    const hv_light* lights;            /* light_count entries */
    float* shadow_matrices;            /* GpuShadowMatrix (float[16] column-major) x matrix_capacity */
    const hv_camera_uniforms* camera;  /* the main camera (CSM cascades) */
    uint32_t* shadow_dirty;            /* one flag per caster slot (shadow_index / 6) */
    uint32_t* shadow_hashes;           /* one FNV-1a hash per caster slot */
    uint32_t light_count;              /* lights processed (the reference's movable light count) */
    uint32_t matrix_capacity;          /* matrices in shadow_matrices */
    uint32_t slot_capacity;            /* entries in shadow_dirty and shadow_hashes */
} hv_shadow_matrix_buffers;

typedef struct hv_shadow_matrix_pass hv_shadow_matrix_pass;

/* ShadowMatrixPass (crates/helio-pass-shadow-matrix/src/lib.rs; shadow_matrices.wgsl:262-315):
 * one lane per light writes 6 cube-face view-projections (point), 4 CSM
 * cascades + 2 identities (directional) or 1 perspective (spot) at
 * shadow_index, hashes the light's 6 matrices (FNV-1a over their bytes) and
 * sets the caster slot's dirty flag when the hash changed. Lights whose
 * matrices or slot fall outside the given capacities are skipped (the
 * reference would address past its buffers). shadow_atlas_size 0 is taken as 1
 * (ShadowMatrixPass::new). tan and acos are this library's canonical
 * binary32 definitions (WGSL leaves their precision to the implementation). */
hv_status hv_shadow_matrix_create(hv_ctx* ctx, uint32_t shadow_atlas_size, hv_shadow_matrix_pass** out);
void hv_shadow_matrix_destroy(hv_shadow_matrix_pass* pass);
/* ShadowMatrixPass::execute: no dispatch when light_count is 0. */
hv_status hv_shadow_matrix_execute(hv_shadow_matrix_pass* pass, const hv_shadow_matrix_buffers* buffers,
                                   void* stream);

/* Plane rule of the per-face sphere test. */
#define HV_SHADOW_PLANES_REFERENCE 0u /* shadow_cull.wgsl:61-73 as written: vp[3] +/- vp[i] are COLUMNS
                                         of the column-major matrix (the planes of its transpose) */
#define HV_SHADOW_PLANES_ROWS 1u      /* Gribb-Hartmann rows, as extract_frustum_planes
                                         (crates/helio-pass-indirect-dispatch/src/lib.rs:271-300) */

typedef struct hv_shadow_cull_buffers {
    const float* shadow_matrices;              /* HV_SHADOW_MAX_FACES float[16] view-projections */
    const hv_instance_data* instances;         /* indexed by draw.first_instance */
    const hv_draw_indexed_indirect* src_indirect; /* the movable shadow draws */
    const uint32_t* face_dirty;                /* HV_SHADOW_MAX_FACES flags; 0 skips the face */
    uint32_t draw_count;                       /* the reference's shadow_movable_draw_count */
    uint32_t face_count;                       /* the reference's shadow_count; 0 = no dispatch */
    uint32_t instance_capacity;                /* a first_instance beyond it reads zeros (radius 0: skipped) */
} hv_shadow_cull_buffers;

typedef struct hv_shadow_cull_pass hv_shadow_cull_pass;

/* ShadowCullPass (crates/helio-pass-shadow-cull/src/lib.rs; shadow_cull.wgsl:75-96):
 * face_counts cleared, then every draw whose instance sphere (r > 0) passes a
 * dirty face's frustum is appended to that face's list
 * (face_indirect[face * max_draws_per_face + slot]); face_counts[face] counts
 * every visible draw, also past max_draws_per_face (those are not written).
 * Slot order inside a face is unspecified, as with the reference's atomic. */
hv_status hv_shadow_cull_create(hv_ctx* ctx, uint32_t max_draws_per_face, uint32_t plane_mode,
                                hv_shadow_cull_pass** out);
void hv_shadow_cull_destroy(hv_shadow_cull_pass* pass);
hv_status hv_shadow_cull_execute(hv_shadow_cull_pass* pass, const hv_shadow_cull_buffers* buffers, void* stream);
/* Device outputs: HV_SHADOW_MAX_FACES * max_draws_per_face commands and
 * HV_SHADOW_MAX_FACES counts (directly usable as multi_draw_indexed_indirect_count
 * count buffers, one per face). */
hv_status hv_shadow_cull_outputs(const hv_shadow_cull_pass* pass, hv_draw_indexed_indirect** face_indirect,
                                 uint32_t** face_counts);
/* Blocking readback of the counts and of each face's written commands
 * (face_indirect: HV_SHADOW_MAX_FACES * max_draws_per_face entries, either may be NULL). */
hv_status hv_shadow_cull_read(hv_shadow_cull_pass* pass, uint32_t counts[HV_SHADOW_MAX_FACES],
                              hv_draw_indexed_indirect* face_indirect);

/* ------------------------------------------------------------------------- */
/* Serving loop: resident scenes, cull requests admitted by device memory,   */
/* continuously batched into graph-replayed launches on several streams.     */
/* No reference counterpart (the reference renders one frame at a time); it  */
/* is built from the batched cull above. Thread-safe.                        */
/* ------------------------------------------------------------------------- */
typedef struct hv_server_config {
    hv_cull_config cull;
    uint32_t max_batch_views;       /* views per launch, 1..1024 (1 = plain per-request path) */
    uint32_t max_inflight;          /* batches on the device at once (one stream each), 1..8 */
    uint64_t max_batch_candidates;  /* candidates per launch (sizes each slot's buffers) */
    uint64_t device_budget_bytes;   /* admission: device bytes of all admitted, unfinished requests */
    uint32_t max_buckets;           /* mesh_count * lod_count per request (sizes args buffers) */
    uint32_t use_graph;             /* replay each slot's captured hipGraph */
    uint32_t copy_ids;              /* copy visible ids back for requests that give ids_out */
    uint32_t max_vg_draws;          /* meshlet draws per launch (sizes each slot's VG outputs);
                                       0: requests with a meshlet part are refused */
    uint32_t grow_pools;            /* 0: each slot's cull pools are allocated for max_batch_candidates at
                                       creation (HV_ENOMEM if they do not fit the device's free memory);
                                       1: pools start empty and grow with the batches admitted, each growth
                                       only while the device's free memory (hipMemGetInfo) stays above
                                       device_headroom_bytes; a batch that would not fit waits, and a head
                                       request that cannot fit even alone, with nothing in flight, fails
                                       with HV_ENOMEM */
    uint64_t device_headroom_bytes; /* device memory left free by pool growth (0: 1 GiB) */
} hv_server_config;

/* Meshlet part of a server request: VirtualGeometryPass semantics
 * (vg_cull.wgsl select + cull) from the request's camera against a resident
 * hv_vg_scene. */
typedef struct hv_vg_request {
    const hv_vg_scene* scene;            /* NULL: no meshlet cull */
    uint32_t width, height;              /* screen: LOD focal length and Hi-Z mip count */
    float lod_error_threshold_px;
    uint32_t flags;                      /* 0 or HV_CULL_REVERSED_Z */
    const hv_hiz* hiz;                   /* optional pyramid (HV_HIZ_MIN with HV_CULL_REVERSED_Z) */
    uint32_t capacity;                   /* publication capacity (max_draw_count), >= 1 */
    hv_draw_indexed_indirect* indirect_out; /* optional host buffer, >= capacity entries */
    hv_vg_draw* metadata_out;            /* optional host buffer, >= capacity entries */
    uint32_t* counters_out;              /* optional host buffer, HV_VG_COUNTER_COUNT entries */
} hv_vg_request;

typedef struct hv_server_request {
    hv_view_request view;               /* the camera is copied at submit; scene may be NULL
                                           (count 0) when the request has a meshlet part */
    uint32_t* ids_out;                  /* optional host buffer, >= view.count entries */
    hv_draw_indexed_indirect* args_out; /* optional host buffer, >= mesh_count * lod_count */
    uint32_t* stats_out;                /* optional host buffer, HV_STAT_COUNT + 1 (visible ids at [8]) */
    hv_vg_request vg;                   /* optional meshlet cull of the same camera */
} hv_server_request;

typedef struct hv_server_stats {
    uint64_t submitted, completed, rejected, batches;
    double latency_p50_us, latency_p99_us, latency_max_us, latency_mean_us;
    double mean_batch_views;
    uint64_t device_bytes_peak;   /* peak admitted-request bytes */
    uint64_t device_bytes_slots;  /* slot buffers held by the server */
    uint64_t device_bytes_free;   /* device free memory at the last admission check */
    uint64_t pool_grows;          /* slot pool growths (grow_pools) */
    uint64_t memory_waits;        /* admission passes cut short by device memory */
} hv_server_stats;

typedef struct hv_server hv_server;

hv_status hv_server_create(hv_ctx* ctx, const hv_server_config* config, hv_server** out);
/* Drains every admitted request, then stops the worker. */
void hv_server_destroy(hv_server* server);
/* Queue one request. HV_ENOMEM if it can never fit the device budget or a
 * slot, HV_EINVAL for a bad request. *ticket identifies it. */
hv_status hv_server_submit(hv_server* server, const hv_server_request* request, uint64_t* ticket);
/* Wait up to timeout_ms (< 0: forever) for a ticket; HV_EAGAIN on time-out.
 * On success the request's output buffers are filled and *visible is set:
 * visible instance ids plus published meshlet draws (min(attempted, capacity)). */
hv_status hv_server_wait(hv_server* server, uint64_t ticket, int32_t timeout_ms, uint32_t* visible);
hv_status hv_server_stats_get(hv_server* server, hv_server_stats* out);
hv_status hv_server_stats_reset(hv_server* server);
/* Open-loop load driver: submits requests[i] at arrival_s[i] seconds after
 * the call (spin-timed on a client thread), waits for all of them and
 * returns the per-request latencies (seconds, submit -> result delivered). */
hv_status hv_server_run_trace(hv_server* server, const hv_server_request* requests, const double* arrival_s,
                              uint32_t count, double* latency_s, double* wall_s);

/* ------------------------------------------------------------------------- */
/* Launch-shape autotuner. Shapes apply to batches/passes created afterwards. */
/* ------------------------------------------------------------------------- */
typedef struct hv_tune_result {
    uint32_t wg;          /* classify workgroup size */
    uint32_t ipt;         /* instances per thread per work item */
    float classify_ms;    /* median classify-kernel time on the tuned batch */
    float total_ms;       /* median of classify + column scan + scatter (what the tuner minimises) */
} hv_tune_result;

hv_status hv_ctx_set_shapes(hv_ctx* ctx, uint32_t wg, uint32_t ipt, uint32_t hiz_tile);
hv_status hv_ctx_get_shapes(hv_ctx* ctx, uint32_t* wg, uint32_t* ipt, uint32_t* hiz_tile);
/* Times every supported (wg, ipt) on the views of a prepared batch (`reps`
 * timed runs each, median of the three kernels' sum), leaves the batch
 * prepared with the fastest and makes it the context default. results may be
 * NULL. */
hv_status hv_autotune_batch(hv_batch* batch, uint32_t reps, hv_tune_result* results, uint32_t capacity,
                            uint32_t* count, hv_tune_result* best);
/* The same, each shape timed as a replay of the batch's captured graph (the
 * three kernels back to back, as a renderer replays them every frame):
 * total_ms is the median from events around the replay, classify_ms is 0. */
hv_status hv_autotune_batch_replay(hv_batch* batch, uint32_t reps, hv_tune_result* results, uint32_t capacity,
                                   uint32_t* count, hv_tune_result* best);
/* Times the pyramid build with 32x32 and 64x64 tiles (ms[0], ms[1]) and makes
 * the faster the context default. */
hv_status hv_autotune_hiz(hv_hiz* hiz, uint32_t reps, float ms[2], uint32_t* best_tile);

#ifdef __cplusplus
}
#endif
#endif /* HV_HV_H */
