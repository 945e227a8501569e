// Internal declarations shared by the host runtime and the gfx950 kernels of
// libhv.so. Not part of the C ABI (include/hv/hv.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "hv/hv.h"

// ---------------------------------------------------------------------------
// Layout contracts of the GPU-shared structs (libhelio), checked at compile time.
// ---------------------------------------------------------------------------
static_assert(sizeof(hv_camera_uniforms) == 368, "GpuCameraUniforms is 368 bytes");
static_assert(offsetof(hv_camera_uniforms, view_proj) == 128, "view_proj @128");
static_assert(offsetof(hv_camera_uniforms, position_near) == 256, "position_near @256");
static_assert(offsetof(hv_camera_uniforms, prev_view_proj) == 304, "prev_view_proj @304");
static_assert(sizeof(hv_instance_data) == 144, "GpuInstanceData is 144 bytes");
static_assert(offsetof(hv_instance_data, bounds) == 112, "bounds @112");
static_assert(offsetof(hv_instance_data, mesh_id) == 128, "mesh_id @128");
static_assert(offsetof(hv_instance_data, flags) == 136, "flags @136");
static_assert(sizeof(hv_instance_aabb) == 32, "GpuInstanceAabb is 32 bytes");
static_assert(sizeof(hv_draw_call) == 20, "GpuDrawCall is 20 bytes");
static_assert(sizeof(hv_draw_indexed_indirect) == 20, "DrawIndexedIndirectArgs is 20 bytes");
static_assert(sizeof(hv_meshlet_entry) == 64, "GpuMeshletEntry is 64 bytes");
static_assert(offsetof(hv_meshlet_entry, first_index) == 48, "first_index @48");
static_assert(sizeof(hv_vg_object) == 128, "GpuVgObject is 128 bytes");
static_assert(offsetof(hv_vg_object, lod_errors) == 32, "lod_errors @32");
static_assert(offsetof(hv_vg_object, lod_first_meshlets) == 64, "lod_first_meshlets @64");
static_assert(offsetof(hv_vg_object, lod_meshlet_counts) == 96, "lod_meshlet_counts @96");
static_assert(sizeof(hv_vg_draw) == 16, "GpuVgDraw is 16 bytes");
static_assert(sizeof(hv_vg_work_item) == 8, "GpuVgWorkItem is 8 bytes");
static_assert(sizeof(hv_instance_cull_data) == 16, "InstanceCullData is 16 bytes");
static_assert(sizeof(hv_id_cull_uniforms) == 112, "indirect-dispatch CullUniforms is 112 bytes");
static_assert(sizeof(hv_occl_cull_params) == 56, "occlusion CullParams is 56 bytes");
static_assert(sizeof(hv_vg_cull_uniforms) == 48, "VG CullUniforms is 48 bytes");
static_assert(sizeof(hv_hiz_uniforms) == 16, "HiZUniforms is 16 bytes");
static_assert(sizeof(hv_mesh_lod) == 12, "hv_mesh_lod is 12 bytes");
static_assert(sizeof(hv_object_info) == 16, "hv_object_info is 16 bytes");
static_assert(sizeof(hv_material_key) == 16, "hv_material_key is 16 bytes");
static_assert(sizeof(hv_material_range) == 24, "hv_material_range is 24 bytes");

namespace hv {

// ---------------------------------------------------------------------------
// Errors: thread-local message, status codes.
// ---------------------------------------------------------------------------
hv_status set_error(hv_status code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
hv_status hip_error(hipError_t e, const char* what, const char* file, int line);

// Workgroups for n items, computed in 64 bits (n near 2^32 must not wrap).
inline uint32_t ceil_div(uint64_t n, uint32_t d) { return (uint32_t)((n + d - 1u) / d); }

#define HV_HIP(call)                                                   \
    do {                                                               \
        hipError_t hv_e_ = (call);                                     \
        if (hv_e_ != hipSuccess) return ::hv::hip_error(hv_e_, #call, __FILE__, __LINE__); \
    } while (0)

#define HV_TRY(call)                      \
    do {                                  \
        hv_status hv_s_ = (call);         \
        if (hv_s_ != HV_OK) return hv_s_; \
    } while (0)

// ---------------------------------------------------------------------------
// Device buffer that grows by doubling (GrowableBuffer counterpart).
// ---------------------------------------------------------------------------
struct DevBuf {
    void* ptr = nullptr;
    uint64_t bytes = 0;
    hv_status reserve(uint64_t need, bool keep_contents, hipStream_t stream);
    void release();
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// ---------------------------------------------------------------------------
// Launch shape of the instance-cull kernels (chosen by the autotuner).
// ---------------------------------------------------------------------------
struct CullShape {
    uint32_t wg = 512;   // threads per workgroup: 64, 128, 256, 512
    uint32_t ipt = 2;    // instances per thread per work item: 1, 2, 4, 8
    uint32_t chunk() const { return wg * ipt; }
};

struct HizShape {
    uint32_t tile = 64;  // mip-0 tile edge reduced by one workgroup: 32 or 64
};

// Looks up this GPU's copy of the no-allocation cull's static scratch array
// (hv_inst_cull_noalloc) once per device; `device` must be current.
hv_status cull_scratch_init(int device);

}  // namespace hv

struct hv_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t cu_count = 0;
    uint64_t hbm_bytes = 0;
    hipEvent_t timers[64] = {};
    hv::CullShape cull_shape;
    hv::HizShape hiz_shape;
    uint32_t occupancy_k1 = 0;   // resident classify workgroups per CU for cull_shape
};

struct hv_scene {
    hv_ctx* ctx = nullptr;
    uint32_t count = 0;          // instances uploaded (high-water mark)
    uint32_t capacity = 0;
    hv::DevBuf bounds;           // float4 per instance (xyz center, w radius)
    hv::DevBuf meta;             // u32: mesh_id | casts_shadow << 31
    hv::DevBuf aabb;             // hv_instance_aabb per instance (explicit mode only)
    hv::DevBuf staging;          // AoS upload staging
    bool explicit_aabb = false;
    hv::DevBuf mesh_lods;        // hv_mesh_lod[mesh_count * lod_count]
    uint32_t mesh_count = 0;
    uint32_t lod_count = 0;
    uint32_t max_mesh_id = 0;    // largest mesh_id ever uploaded (host-side check)
    // Bumped whenever the device arrays move or change shape (growth, the switch
    // to explicit AABBs and back, a new LOD table): passes and batches prepared
    // before it refuse to execute on stale pointers.
    uint64_t storage_gen = 0;
};

// Grows the resident arrays to hold `need` instances (x2 steps, contents kept).
hv_status scene_grow(hv_scene* s, uint32_t need, hipStream_t stream);

struct hv_hiz {
    hv_ctx* ctx = nullptr;
    uint32_t width = 0, height = 0, levels = 0, mode = 0;
    hv::DevBuf own_level0;
    const float* level0 = nullptr;   // bound depth buffer or own_level0
    hv::DevBuf packed;               // levels 1..levels-1
    uint64_t level_off[16] = {};     // float offsets of level m >= 1 inside `packed`
    uint64_t last_generation = ~0ull;
    bool skip_next = false;
    uint64_t storage_gen = 0;        // bumped when level0 points elsewhere (prepared passes check it)
};
