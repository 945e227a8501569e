/*
 * TEST INFRASTRUCTURE ONLY — the CPU oracle of the hv visibility path.
 *
 * A plain C++ restatement of the reference's WGSL compute shaders (t4ce/Helio)
 * and of the host-side math around them. It is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it. Nothing in the product (helio_amd/, include/) links or calls it.
 *
 * Float semantics (the canonical definition shared with the HIP kernels):
 * IEEE binary32, compiled with -ffp-contract=off (no FMA), correctly rounded
 * `/` and sqrt, denormals preserved; dot products and matrix x vector are
 * summed left to right ((x + y) + z) + w; WGSL max(a,b) = (a < b) ? b : a,
 * min(a,b) = (b < a) ? b : a, clamp(e,lo,hi) = min(max(e,lo),hi);
 * ceil(log2(d)) for d >= 1 is computed exactly from the exponent and mantissa;
 * nearest texel = clamp(floor(u * w_m), 0, w_m - 1).
 */
#ifndef HV_ORACLE_H
#define HV_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One view of the per-instance compacting cull. */
typedef struct or_view {
    float planes[24];      /* from or_extract_frustum_planes(view_proj) */
    float view_proj[16];
    float p00, p11;        /* proj[0][0], proj[1][1] */
    float cam[3];          /* camera position */
    float lod_distances[7];
    uint32_t lod_count;
    uint32_t flags;        /* 1 = occlusion, 2 = reversed Z */
    uint32_t screen_height;
    uint32_t hiz_mip_count;   /* CullParams.hiz_mip_count = mip_levels(screen) */
    uint32_t mesh_count;
    const uint32_t* mesh_lods; /* mesh_count*lod_count x {index_count, first_index, base_vertex} */
} or_view;

/* A Hi-Z pyramid: level 0 plus levels 1.. packed row-major one after another. */
typedef struct or_pyramid {
    const float* level0;
    const float* packed;   /* level 1 first */
    uint32_t width, height, levels;
    uint32_t mode;         /* 0 = max (standard Z), 1 = min (reversed Z) */
} or_pyramid;

uint32_t or_mip_levels(uint32_t w, uint32_t h);
void or_extract_frustum_planes(const float vp[16], float planes[24]);
uint32_t or_ceil_log2(float d);
uint32_t or_pick_mip(float r_px, uint32_t mip_count);
int or_is_occluded(float near_z, float hiz_depth);
uint64_t or_hiz_packed_floats(uint32_t w, uint32_t h);
void or_hiz_build(const float* level0, uint32_t w, uint32_t h, uint32_t mode, float* packed);

/* Per-instance classification. code[i] = bucket (mesh*lod_count + lod) for a
 * visible instance, 0xFFFFFFFF otherwise; stats accumulate (8 slots). */
void or_classify(const or_view* v, const float* bounds4, const uint32_t* meta,
                 const float* aabb6 /* NULL => sphere_to_aabb */, uint32_t first, uint32_t count,
                 const or_pyramid* pyr, uint32_t* code, uint32_t stats[8]);

/* Full compacting cull of one view with `threads` host threads. ids are
 * bucket-major, ascending instance id inside a bucket. args has
 * mesh_count*lod_count entries (5 u32 each). stats: 8 slots. Returns the
 * number of ids. */
uint32_t or_cull_view(const or_view* v, const float* bounds4, const uint32_t* meta,
                      const float* aabb6, uint32_t first, uint32_t count, const or_pyramid* pyr,
                      uint32_t* ids, uint32_t* args, uint32_t stats[8], uint32_t threads);

/* --- reference per-draw passes ------------------------------------------ */
/* indirect_dispatch.wgsl:100-176 (instances: 144 B AoS, draws: 20 B, aabbs: 32 B). */
void or_indirect_dispatch(const float camera[92], const float planes[24], const void* instances,
                          const void* draw_calls, const void* aabbs, uint32_t draw_count,
                          uint32_t* indirect /* 5 u32 per draw */, uint32_t stats[8]);
/* occlusion_cull.wgsl:146-220 (static-Hi-Z branch excluded). */
void or_occlusion_cull_draws(const float camera[92], uint32_t screen_w, uint32_t screen_h,
                             uint32_t hiz_mip_count, const void* instances, const void* draw_calls,
                             uint32_t draw_count, const or_pyramid* pyr, uint32_t* indirect,
                             uint32_t stats[8]);

/* --- virtual geometry ----------------------------------------------------- */
void or_instance_cull_data(const void* instance /* 144 B */, float* max_scale, float* min_scale,
                           uint32_t* cone_enabled, uint32_t* valid);
uint32_t or_select_object_lod(const float errors[8], uint32_t lod_count, float max_scale,
                              float focal_pixels, float closest_distance, float max_error_pixels);
/* vg_cull.wgsl cs_select_objects (:282-352) then cs_cull_meshlets (:354-405)
 * executed in work-item order, lanes in order, so slots are assigned
 * sequentially. objects is modified (selected_lod_plus_one). */
typedef struct or_vg_inputs {
    const float* camera;            /* 92 floats (hv_camera_uniforms) */
    const uint32_t* uniforms;       /* 12 u32 (hv_vg_cull_uniforms) */
    const void* meshlets; uint32_t meshlet_count;
    void* objects; uint32_t object_count_buf;
    const void* instances; uint32_t instance_count;
    const void* instance_cull; uint32_t instance_cull_count;
    const void* work_items; uint32_t work_item_buf_count;
    uint32_t indirect_len, metadata_len;  /* buffer lengths (entries) */
    uint32_t counter_len;                 /* 2 or 11 */
    const or_pyramid* pyr;
    uint32_t reversed_z;
} or_vg_inputs;
void or_vg_cull(const or_vg_inputs* in, uint32_t* indirect /* 5 u32 per slot */,
                uint32_t* metadata /* 4 u32 per slot */, uint32_t* counters);

/* --- multi-view producers ---------------------------------------------------
 * Canonical binary32 tan and acos of the library (WGSL leaves their precision
 * to the implementation): polynomial fits evaluated with + - * / and sqrt in a
 * fixed order, so the oracle and the HIP kernels agree bit for bit. */
float or_tanf(float x);
float or_acosf(float x);
/* FNV-1a (32-bit) of one matrix's 64 bytes in memory order (shadow_matrices.wgsl:238-252). */
uint32_t or_fnv_mat(const float m[16]);
/* compute_shadow_matrices (shadow_matrices.wgsl:262-315) for lights 0..n-1 in
 * order. lights: 64-byte GpuLight records; camera: 92 floats; mats:
 * matrix_capacity float[16]; dirty/hashes: slot_capacity u32. Lights whose six
 * matrices or caster slot fall outside the capacities are skipped. */
void or_shadow_matrices(const void* lights, uint32_t light_count, const float* camera, uint32_t atlas_size,
                        float* mats, uint32_t matrix_capacity, uint32_t* dirty, uint32_t* hashes,
                        uint32_t slot_capacity);
/* shadow_cull.wgsl main (:75-96) over draws 0..n-1 in order: counts[256]
 * cleared first, visible draws appended per face in draw order to
 * dst[(face * max_per_face + slot) * 5]. plane_mode 0 = the shader's column
 * vectors, 1 = Gribb-Hartmann rows. A first_instance >= instance_capacity
 * reads a zero sphere. */
void or_shadow_cull(const float* mats, const void* instances, uint32_t instance_capacity,
                    const uint32_t* src_indirect, uint32_t draw_count, const uint32_t* face_dirty,
                    uint32_t max_per_face, uint32_t plane_mode, uint32_t* counts, uint32_t* dst);

/* --- instance/draw rebuild ----------------------------------------------------
 * Scene::rebuild_instance_buffers_persistent (rebuild.rs:91-191, mode 0) or
 * _optimized (rebuild.rs:234-361, mode 1), then rebuild_shadow_partition_buffers
 * (rebuild.rs:374-425). instances: n 144-byte GpuInstanceData; aabbs: n 32-byte
 * records; draws: n GpuDrawCall (5 u32); objects: n hv_object_info (u64 groups,
 * u32 material_slot, u32 movability); materials: m_count (u64 graph_hash, u32
 * class, u32 pad). Outputs hold n entries each; ranges are 24-byte (u64
 * graph_hash, u32 class, start, count, pad). counts = [draws, ranges,
 * shadow static, shadow movable]; with n == 0 only counts[0..1] are written. */
void or_scene_rebuild(uint32_t mode, uint32_t n, const void* instances, const void* aabbs, const uint32_t* draws,
                      const void* objects, const void* materials, uint32_t m_count, uint64_t group_hidden,
                      void* out_instances, void* out_aabbs, uint32_t* out_draw_calls, uint32_t* out_indirect,
                      uint32_t* visibility, uint32_t* gpu_slots, void* ranges, uint32_t* shadow_static,
                      uint32_t* shadow_movable, uint32_t* counts);

#ifdef __cplusplus
}
#endif
#endif
